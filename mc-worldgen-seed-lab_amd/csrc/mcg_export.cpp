// mcg_export.cpp - an engine's vanilla world for one seed in libmcgen GPU-plugin form (McgBiomeWorld).
//
// libmcgen builds these descriptions from datapack density functions (libmcgen/src/gpu_bridge.c:246-315,
// gpu_export_new.inc, gpu_export_old.inc).  The engine carries the vanilla climate in its own tables,
// so it can write the same router directly: the climate sampler's computation as a node program in
// libmcgen's lowering (shift_a = CMUL(NOISE(0.25, 0), 4), shift_b = SHIFT_B, shifted_noise =
// NOISE_XZ / SHIFTED_NOISE, constant operands folded into CADD / CMUL, caches as pass-through nodes,
// blend_alpha / blend_offset as context constants), noise stacks in libmcgen's layer order, the seed's
// octave tables from the device seed-state kernel, and the biome tree laid out breadth-first with
// consecutive children.  The offset (7th) dimension of a tree node is written as the degenerate interval
// whose squared distance to 0 is the node's d6, which is all a search reads of it.
// Used to drive mcgpu_biome_* with vanilla worlds without libmcgen (tests, benchmarks, embedding).
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/mcgpu.h"
#include "mcb_internal.h"
#include "mcg_plugin.h"

using namespace mcb;

namespace {

const double kInputFactor = 1.0181268882175227;

struct Owned {
    McgBiomeWorld W;                      // first member: the caller's pointer
    McgProg P;
    std::vector<McgNode> nodes;
    std::vector<McgNoise> noise;
    std::vector<McgLayer> layers;
    std::vector<McgOct> octs;
    std::vector<McgSpline> sp;
    std::vector<float> splf;
    std::vector<int> spc, arr;
    std::vector<double> thr;
    std::vector<McgTreeNode> tree;
};
struct OwnedTerrain {
    McgTerrainWorld T;                    // first member: the caller's pointer
    Owned prog;
};

McgNode node(int k) {
    McgNode n;
    memset(&n, 0, sizeof n);
    n.k = k;
    n.a = n.b = n.c = -1;
    n.ns = n.sp = n.oct = -1;
    return n;
}

struct Builder {
    Owned &o;
    const McbWorld &w;
    const McbSlot &slot;
    int add(McgNode n) { o.nodes.push_back(n); return (int)o.nodes.size() - 1; }
    int un(int k, int a) { McgNode n = node(k); n.a = a; return add(n); }
    int unf(int k, int a, float f0, double d0) { McgNode n = node(k); n.a = a; n.f0 = f0; n.d0 = d0; return add(n); }
    int oct(int o_) {
        McgOct m;
        m.xo = slot.oct[o_].xo; m.yo = slot.oct[o_].yo; m.zo = slot.oct[o_].zo;
        memcpy(m.p, slot.perm[o_], 256);
        o.octs.push_back(m);
        return (int)o.octs.size() - 1;
    }
    // NormalNoise i: 26.3 one stack (levels' two octaves interleaved; legacy Nether form: first stack then
    // second), 26.1/26.2 two PerlinNoises (libmcgen/src/noise.c:453-500, 578-603)
    int noise(int i) {
        const McbNoise &N = w.noise[i];
        const int n = N.nlev, base = w.oct_base[i];
        McgNoise mn;
        memset(&mn, 0, sizeof mn);
        if (w.fmt_new) {
            mn.l0 = (int)o.layers.size();
            mn.n0 = 2 * n;
            for (int q = 0; q < 2 * n; q++) {
                const int stack = N.legacy ? q / n : (q & 1), k = N.legacy ? q % n : (q >> 1);
                McgLayer L;
                memset(&L, 0, sizeof L);
                L.a = stack ? N.freq[k] * kInputFactor : N.freq[k];
                L.amp_f = N.amp_f[k];
                L.kind = 0;
                L.oct = oct(base + stack * n + k);
                o.layers.push_back(L);
            }
        } else {
            for (int stack = 0; stack < 2; stack++) {
                (stack ? mn.l1 : mn.l0) = (int)o.layers.size();
                for (int k = 0; k < n; k++) {
                    McgLayer L;
                    memset(&L, 0, sizeof L);
                    L.a = N.freq[k];
                    L.b = N.vf_d[k];
                    L.amp_d = N.amp_d[k];
                    L.oct = oct(base + stack * n + k);
                    o.layers.push_back(L);
                }
                (stack ? mn.n1 : mn.n0) = n;
            }
            mn.value_factor = N.value_factor;
        }
        o.noise.push_back(mn);
        return (int)o.noise.size() - 1;
    }
    void spline() {
        const McbSpline &S = w.spline;
        for (int id = 0; id < S.n_nodes; id++) {
            McgSpline m;
            memset(&m, 0, sizeof m);
            if (S.kind[id] == 0) {
                m.is_const = 1;
                m.value = S.value[id];
            } else {
                m.coord = S.coord[id];
                m.n = S.count[id];
                m.loc0 = (int)o.splf.size();
                for (int i = 0; i < S.count[id]; i++) o.splf.push_back(S.loc[S.first[id] + i]);
                for (int i = 0; i < S.count[id]; i++) o.splf.push_back(S.der[S.first[id] + i]);
                m.val0 = (int)o.spc.size();
                for (int i = 0; i < S.count[id]; i++) o.spc.push_back(S.child[S.first[id] + i]);
            }
            o.sp.push_back(m);
        }
    }
};

// pre-order array with escape links -> breadth-first array with consecutive children
void tree_bfs(const std::vector<McbNode> &T, std::vector<McgTreeNode> &out) {
    out.clear();
    if (T.empty()) return;
    std::vector<int> order{0};
    out.resize(T.size());
    size_t head = 0;
    while (head < order.size()) {
        const int i = order[head];
        const McbNode &nd = T[i];
        McgTreeNode &m = out[head];
        memset(&m, 0, sizeof m);
        for (int k = 0; k < 6; k++) { m.lo[k] = nd.lo[k]; m.hi[k] = nd.hi[k]; }
        const int32_t x = (int32_t)llround(std::sqrt((double)nd.d6));
        m.lo[6] = m.hi[6] = x;
        if (nd.leaf & 0x8000) {
            m.count = 0;
            m.first = -1;
            m.biome = nd.leaf & 0xFF;
        } else {
            m.biome = -1;
            m.first = (int32_t)order.size();
            for (int c = i + 1; c < (int)nd.escape; c = T[c].escape) order.push_back(c);
            m.count = (int32_t)order.size() - m.first;
        }
        head++;
    }
    out.resize(order.size());
}

// The vanilla Overworld climate router: shift_x / shift_z, the five shifted climate noises, folded ridges,
// the offset spline and depth, in libmcgen's lowering of the 26.3+ program (float) or the legacy one.
struct Router {
    int ns[6];                    // McgNoise of the engine's six noises
    int f[5];                     // temperature, vegetation, continents, erosion, weirdness (ridges)
    int offset = -1;              // 26.3+: the blended offset (y-independent)
    int depth = -1;
};
Router overworld_router(Builder &B, const McbWorld &w, bool N) {
    Router R;
    int *ns = R.ns;
    for (int i = 0; i < 6; i++) ns[i] = B.noise(i);
    int sx, sz, f[5];
    if (N) {
        McgNode a = node(MCGK_NOISE); a.ns = ns[0]; a.d0 = 0.25; a.d1 = 0.0;
        sx = B.unf(MCGK_CMUL, B.add(a), 4.0f, 0.0);
        McgNode b = node(MCGK_SHIFT_B); b.ns = ns[0]; b.d0 = 0.25; b.d1 = 0.25;
        sz = B.add(b);
        sx = B.un(MCGK_CACHE, sx);
        sz = B.un(MCGK_CACHE, sz);
        for (int i = 0; i < 5; i++) {
            McgNode c = node(MCGK_NOISE_XZ); c.ns = ns[1 + i]; c.d0 = 0.25; c.d1 = 0.0; c.a = sx; c.c = sz;
            f[i] = B.add(c);
        }
    } else {
        McgNode a = node(MCGO_SHIFT_A); a.ns = ns[0]; a.d0 = 0.25;
        McgNode b = node(MCGO_SHIFT_B); b.ns = ns[0]; b.d0 = 0.25;
        sx = B.un(MCGO_PASS, B.add(a));
        sz = B.un(MCGO_PASS, B.add(b));
        McgNode z = node(MCGO_CONST); z.d0 = 0.0;
        const int zero = B.add(z);
        for (int i = 0; i < 5; i++) {
            McgNode c = node(MCGO_SHIFTED_NOISE); c.ns = ns[1 + i]; c.d0 = 0.25; c.d1 = 0.0; c.a = sx; c.b = zero; c.c = sz;
            f[i] = B.add(c);
        }
    }
    const int C = f[2], E = f[3], Wr = f[4];
    int rf;                                                       // (|(|W| + r0)| + r1) * r2
    if (N)
        rf = B.unf(MCGK_CMUL, B.unf(MCGK_CADD, B.un(MCGK_ABS, B.unf(MCGK_CADD, B.un(MCGK_ABS, Wr), w.rf_f[0], 0.0)), w.rf_f[1], 0.0),
                   w.rf_f[2], 0.0);
    else
        rf = B.unf(MCGO_CMUL, B.unf(MCGO_CADD, B.un(MCGO_ABS, B.unf(MCGO_CADD, B.un(MCGO_ABS, Wr), 0.0f, w.rf_d[0])), 0.0f, w.rf_d[1]),
                   0.0f, w.rf_d[2]);
    B.spline();
    McgNode s = node(N ? MCGK_SPLINE : MCGO_SPLINE);
    s.sp = 0;
    s.arr0 = (int)B.o.arr.size();
    s.narr = 4;
    B.o.arr.insert(B.o.arr.end(), {C, E, Wr, rf});
    const int spl = B.add(s);
    int D;
    if (N) {
        const int off = B.unf(MCGK_CADD, spl, w.offc_f, 0.0);
        McgNode al = node(MCGK_CTX_ALPHA); al.f0 = 1.0f;
        McgNode bo = node(MCGK_CTX_OFFSET); bo.f0 = 0.0f;
        McgNode le = node(MCGK_LERP); le.a = B.add(al); le.b = B.add(bo); le.c = off;
        const int blended = B.un(MCGK_CACHE, B.add(le));
        R.offset = blended;
        McgNode g = node(MCGK_GRAD_CLAMP);
        g.i0 = 1; g.i1 = w.gy0; g.i2 = std::min(w.gy0, w.gy1); g.i3 = std::max(w.gy0, w.gy1);
        g.f0 = w.gv0_f; g.f1 = (w.gv1_f - w.gv0_f) / (float)(w.gy1 - w.gy0);
        McgNode ad = node(MCGK_ADD); ad.a = B.add(g); ad.b = blended;
        D = B.add(ad);
    } else {
        const int off = B.unf(MCGO_CADD, B.unf(MCGO_CMUL, B.unf(MCGO_CADD, spl, 0.0f, w.offc_d), 0.0f, 1.0), 0.0f, 0.0);
        McgNode g = node(MCGO_Y_GRADIENT); g.i0 = w.gy0; g.i1 = w.gy1; g.d0 = w.gv0_d; g.d1 = w.gv1_d;
        McgNode ad = node(MCGO_ADD); ad.a = B.add(g); ad.b = off;
        D = B.add(ad);
    }
    for (int i = 0; i < 5; i++) R.f[i] = f[i];
    R.depth = D;
    return R;
}

// the seed's octave state, from the device seed-state kernel
int seed_slot(McbEngine *e, int64_t seed, McbSlot &slot) {
    std::lock_guard<std::mutex> lk(e->mu);
    if (int rc = dev_grow(&e->d_slots, &e->slots_cap, (size_t)1)) return rc;
    if (int rc = dev_grow(&e->d_seeds, &e->seeds_cap, (size_t)1)) return rc;
    if (int rc = ensure_ident(e, 1)) return rc;
    MCB_HIPCHK(hipMemcpy(e->d_seeds, &seed, sizeof seed, hipMemcpyHostToDevice));
    MCB_HIPCHK(launch_seed_init(e->w, e->d_seeds, e->d_ident, 1, e->d_slots, nullptr));
    MCB_HIPCHK(hipMemcpy(&slot, e->d_slots, sizeof slot, hipMemcpyDeviceToHost));
    return MCB_OK;
}

void finish_prog(Owned *o, bool old, const int *root, int nroots) {
    McgProg &P = o->P;
    P.old = old;
    P.nnodes = (int)o->nodes.size(); P.nodes = o->nodes.data();
    P.nnoise = (int)o->noise.size(); P.noise = o->noise.data();
    P.nlayers = (int)o->layers.size(); P.layers = o->layers.data();
    P.noct = (int)o->octs.size(); P.octs = o->octs.data();
    P.nsp = (int)o->sp.size(); P.sp = o->sp.data();
    P.nsplf = (int)o->splf.size(); P.splf = o->splf.data();
    P.nspc = (int)o->spc.size(); P.spc = o->spc.data();
    P.narr = (int)o->arr.size(); P.arr = o->arr.data();
    P.nthr = 0; P.thr = nullptr;
    P.nroots = nroots;
    for (int k = 0; k < nroots; k++) P.root[k] = root[k];
}

}  // namespace

extern "C" {

int mcb_plugin_world(McbEngine *e, int64_t seed, McgBiomeWorld **out) {
    if (!e || !out) return fail(MCB_E_ARG, "bad arguments");
    *out = nullptr;
    // the plugin ABI's tree nodes carry closed intervals (libmcgen/gpu/mcgpu_dev.cuh:463-473)
    if (e->w.version >= MCB_V26_4) return fail(MCB_E_ARG, "26.4-snapshot-2 trees (exclusive upper bounds) have no plugin form");
    const McbWorld &w = e->w.host;
    McbSlot slot;
    if (int rc = seed_slot(e, seed, slot)) return rc;
    Owned *o = new Owned();
    memset(&o->W, 0, sizeof o->W);
    memset(&o->P, 0, sizeof o->P);
    Builder B{*o, w, slot};
    McgBiomeWorld &W = o->W;
    const bool N = w.fmt_new != 0;
    W.newf = N;
    W.zoom_seed = hashed_seed(seed);
    for (int i = 0; i < 5; i++) W.end_ids[i] = w.end_ids[i];
    int root[6] = {-1, -1, -1, -1, -1, -1};
    if (w.dim == MCB_OVERWORLD) {
        W.source = MCG_BS_MULTI; W.min_qy = -16; W.qh = 96;
        const Router R = overworld_router(B, w, N);
        const int T = R.f[0], H = R.f[1], C = R.f[2], E = R.f[3], D = R.depth, Wr = R.f[4];
        root[0] = T; root[1] = H; root[2] = C; root[3] = E; root[4] = D; root[5] = Wr;
        tree_bfs(e->tree_host, o->tree);
    } else if (w.dim == MCB_NETHER) {
        W.source = MCG_BS_MULTI; W.min_qy = 0; W.qh = 64;
        const int nt = B.noise(0), nv = B.noise(1);
        int zero;
        if (N) {
            McgNode t = node(MCGK_NOISE); t.ns = nt; t.d0 = 0.25; t.d1 = 0.0; root[0] = B.add(t);
            McgNode v = node(MCGK_NOISE); v.ns = nv; v.d0 = 0.25; v.d1 = 0.0; root[1] = B.add(v);
            McgNode z = node(MCGK_CONST); z.f0 = 0.0f; zero = B.add(z);
        } else {
            McgNode z = node(MCGO_CONST); z.d0 = 0.0; zero = B.add(z);
            McgNode t = node(MCGO_SHIFTED_NOISE); t.ns = nt; t.d0 = 0.25; t.d1 = 0.0; t.a = t.b = t.c = zero; root[0] = B.add(t);
            McgNode v = node(MCGO_SHIFTED_NOISE); v.ns = nv; v.d0 = 0.25; v.d1 = 0.0; v.a = v.b = v.c = zero; root[1] = B.add(v);
        }
        for (int k = 2; k < 6; k++) root[k] = zero;
        tree_bfs(e->tree_host, o->tree);
    } else {
        W.source = MCG_BS_END; W.min_qy = 0; W.qh = 64;
        if (N) {
            // erosion = max((clamp(100 - |(x, z)|, -100, 80) - 8) / 128, (height - 8) / 128)
            McgNode d = node(MCGK_DIST); d.cm = 2; d.cy = 0;
            const int len = B.add(d);
            const int c0 = B.unf(MCGK_CSUB, len, 100.0f, 0.0);
            McgNode cl = node(MCGK_CLAMP); cl.a = c0; cl.f0 = -100.0f; cl.f1 = 80.0f;
            const int c = B.unf(MCGK_CMUL, B.unf(MCGK_CADD, B.add(cl), -8.0f, 0.0), 0.0078125f, 0.0);
            McgNode en = node(MCGK_END); en.oct = B.oct(0);
            McgNode mx = node(MCGK_MAX); mx.a = c; mx.b = B.add(en); mx.f0 = INFINITY;
            root[3] = B.add(mx);
        } else {
            McgNode en = node(MCGO_END_ISLANDS); en.oct = B.oct(0);
            root[3] = B.add(en);
        }
    }
    finish_prog(o, !N, root, 6);
    W.prog = &o->P;
    W.ntree = (int)o->tree.size();
    W.tree = o->tree.empty() ? nullptr : o->tree.data();
    W.tree_root = 0;
    *out = &o->W;
    return MCB_OK;
}

void mcb_plugin_world_free(McgBiomeWorld *W) { delete reinterpret_cast<Owned *>(W); }

// A terrain world over the engine's vanilla 26.3 Overworld climate, shaped like the game's Overworld
// final_density (whose datapack graph is not part of the engine): the router's erosion and offset read
// through slices at y = 0 (where libmcgen's compiler places 2D functions in a 3D one), depth = y gradient +
// offset, sloped = 4 * quarter_negative(depth * clamp(3.5 - 2 * erosion, 1, 6)) + a 3D noise (the ridge
// noise at xz 0.25, y 0.125), a cave noise (the erosion noise at xz 1, y 2), a top slide, interpolation
// over 4 x 8 x 4 cells and a final squeeze(0.64 * x).  For benchmarks and tests of the terrain executor.
int mcb_plugin_terrain_world(McbEngine *e, int64_t seed, McgTerrainWorld **out) {
    if (!e || !out) return fail(MCB_E_ARG, "bad arguments");
    *out = nullptr;
    const McbWorld &w = e->w.host;
    if (w.dim != MCB_OVERWORLD || !w.fmt_new) return fail(MCB_E_ARG, "terrain worlds are built for 26.3+ Overworld engines");
    McbSlot slot;
    if (int rc = seed_slot(e, seed, slot)) return rc;
    OwnedTerrain *t = new OwnedTerrain();
    Owned *o = &t->prog;
    memset(&o->W, 0, sizeof o->W);
    memset(&o->P, 0, sizeof o->P);
    Builder B{*o, w, slot};
    const Router R = overworld_router(B, w, true);
    auto slice_y = [&](int a) { McgNode n = node(MCGK_SLICE_Y); n.a = a; n.i0 = 0; return B.add(n); };
    auto grad_y = [&](int y0, int y1, float v0, float v1) {
        McgNode g = node(MCGK_GRAD_CLAMP);
        g.i0 = 1; g.i1 = y0; g.i2 = std::min(y0, y1); g.i3 = std::max(y0, y1);
        g.f0 = v0; g.f1 = (v1 - v0) / (float)(y1 - y0);
        return B.add(g);
    };
    auto bin = [&](int k, int a, int b) { McgNode n = node(k); n.a = a; n.b = b; return B.add(n); };
    auto noise3 = [&](int ns, double xz, double y) { McgNode n = node(MCGK_NOISE); n.ns = ns; n.d0 = xz; n.d1 = y; return B.add(n); };
    const int erosion = slice_y(R.f[3]), offset = slice_y(R.offset);
    const int depth = bin(MCGK_ADD, grad_y(w.gy0, w.gy1, w.gv0_f, w.gv1_f), offset);
    McgNode cl = node(MCGK_CLAMP);
    cl.a = B.unf(MCGK_CADD, B.unf(MCGK_CMUL, erosion, -2.0f, 0.0), 3.5f, 0.0);
    cl.f0 = 1.0f; cl.f1 = 6.0f;
    const int factor = B.add(cl);
    const int sloped = bin(MCGK_ADD, B.unf(MCGK_CMUL, B.unf(MCGK_LEAKY, bin(MCGK_MUL, depth, factor), 0.25f, 0.0), 4.0f, 0.0),
                           noise3(R.ns[5], 0.25, 0.125));
    const int caves = B.unf(MCGK_CADD, B.unf(MCGK_CMUL, noise3(R.ns[4], 1.0, 2.0), 5.0f, 0.0), 0.2f, 0.0);
    McgNode mn = node(MCGK_MIN); mn.a = sloped; mn.b = caves; mn.f0 = -INFINITY;
    const int body = bin(MCGK_ADD, grad_y(240, 256, 0.0f, -0.078125f), B.add(mn));
    McgNode ip = node(MCGK_INTERP);
    ip.a = B.un(MCGK_BLEND_DENSITY, body);
    ip.i0 = 4; ip.i1 = 8; ip.f0 = 0.25f; ip.f1 = 0.125f;
    const int root = B.un(MCGK_SQUEEZE, B.unf(MCGK_CMUL, B.add(ip), 0.64f, 0.0));
    finish_prog(o, false, &root, 1);
    McgTerrainWorld &T = t->T;
    memset(&T, 0, sizeof T);
    T.vol = &o->P;
    T.nmin = w.gy0;
    T.nh = w.gy1 - w.gy0;
    T.pre_root = -1;
    *out = &t->T;
    return MCB_OK;
}

void mcb_plugin_terrain_world_free(McgTerrainWorld *T) { delete reinterpret_cast<OwnedTerrain *>(T); }

}  // extern "C"

// Renumber the biome ids of an exported world (tree leaves, End ids, fixed biome) through map[old id].
void mcb::plugin_world_renumber(McgBiomeWorld *W, const uint8_t map[256]) {
    Owned *o = reinterpret_cast<Owned *>(W);
    for (auto &t : o->tree)
        if (t.count == 0 && t.biome >= 0 && t.biome < 256) t.biome = map[t.biome];
    for (int k = 0; k < 5; k++)
        if (W->end_ids[k] >= 0 && W->end_ids[k] < 256) W->end_ids[k] = map[W->end_ids[k]];
    if (W->fixed_biome >= 0 && W->fixed_biome < 256) W->fixed_biome = map[W->fixed_biome];
}

extern "C" {

// Whole chunks as libmcgen stores them (world_chunk_biomes, libmcgen/src/biome.c:125-157), through the
// plugin kernels on the engine's vanilla world: per cell climate + search, ties resolved along the chunk's
// fill order from its first cell.  The 26.3 chunk path samples the unshifted noises by volume, which sums
// each octave's gradient terms as (x + z) + y instead of (x + y) + z and scales coordinates as
// block * (0.25 * frequency); every gradient vector has a zero component and 0.25 is a power of two, so
// both give the point path's values bit for bit and the point evaluation serves here.
int mcb_chunk_biomes(McbEngine *e, int64_t seed, int n, const int32_t *cxz, uint8_t *out) {
    if (!e || n < 0 || (n && (!cxz || !out))) return fail(MCB_E_ARG, "bad arguments");
    McgBiomeWorld *W = nullptr;
    if (int rc = mcb_plugin_world(e, seed, &W)) return rc;
    char err[512] = {0};
    int bad = 0;
    const int r = mcb::plug::chunk_biomes(W, n, cxz, out, &bad, err, sizeof err);
    mcb_plugin_world_free(W);
    if (r) return fail(MCB_E_DEVICE, "%s", err);
    if (bad) return fail(MCB_E_DATA, "%d cells have a non-finite climate", bad);
    return MCB_OK;
}

}  // extern "C"
