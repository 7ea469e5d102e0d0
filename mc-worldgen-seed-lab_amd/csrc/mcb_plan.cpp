// mcb_plan.cpp - batched plans (one persistent launch per world over many uneven requests), hipGraph
// capture, the tile autotuner, contexts of several worlds, and the plain per-call region path.
//
// A plan fixes everything a run needs on the device: unique seeds and their state slots, the kernel
// request/tile descriptors, the zoom staging buffer.  A run is then only kernel launches (seed state,
// quart-grid tiles, zoomed quart layers, zoom pass) on one stream, so it captures into a hipGraph.
// Replaces the per-call pattern malloc/launch/sync/copy of libmcgen/gpu/mcgpu_biome.cu:203-249.
#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "mcb_internal.h"

using namespace mcb;

namespace mcb {

int check_region(const McbRegion &r) {
    if (r.nx <= 0 || r.nz <= 0 || r.ny <= 0) return fail(MCB_E_ARG, "empty region");
    if (r.ny > (r.zoom ? 4 : 8)) return fail(MCB_E_LIMIT, "too many layers (%d)", r.ny);
    if (r.nx > 32767 * 8 || r.nz > 32767 * 8) return fail(MCB_E_LIMIT, "region too large");
    if (r.out_off < 0) return fail(MCB_E_ARG, "negative out_off");
    return 0;
}

bool wrap_free(const McbWorld &w, const std::vector<McbRequest> &reqs) {
    double fmax = 0.0, shift = 0.0;
    for (int i = 0; i < w.n_noise; i++)
        for (int k = 0; k < w.noise[i].nlev; k++) {
            fmax = std::max(fmax, w.noise[i].freq[k] * 1.0181268882175227 * 1.0181268882175227);
            if (w.dim != MCB_OVERWORLD || i != 0) continue;
            // the offset noise's value: 26.3 sums amp_f * octave over both stacks; 26.1/26.2 sum
            // amp_d * octave * vf_d per stack, then (v1 + v2) * valueFactor
            const McbNoise &N = w.noise[0];
            const double amp = w.fmt_new ? std::fabs((double)N.amp_f[k]) : std::fabs(N.amp_d[k] * N.vf_d[k] * N.value_factor);
            shift += 4.0 * 2.0 * 2.0 * amp;
        }
    double q = 0.0;
    for (const McbRequest &r : reqs) {
        q = std::max(q, std::fabs((double)r.x0));
        q = std::max(q, std::fabs((double)r.x0 + r.nx - 1));
        q = std::max(q, std::fabs((double)r.z0));
        q = std::max(q, std::fabs((double)r.z0 + r.nz - 1));
    }
    return (q + shift + 1.0) * fmax < 8388608.0;
}

static void add_tiles(std::vector<McbTile> &tiles, int req, int nx, int nz, int tw, int th) {
    for (int tz = 0; tz < (nz + th - 1) / th; tz++)
        for (int tx = 0; tx < (nx + tw - 1) / tw; tx++) tiles.push_back(McbTile{req, (int16_t)tx, (int16_t)tz});
}

void lower_region(const McbRegion &r, int slot, int tile_w, int cpl, Lowered &L) {
    const int tile_h = 256 / tile_w * cpl;
    McbRequest R;
    memset(&R, 0, sizeof R);
    R.slot = slot;
    R.world = r.world;
    if (!r.zoom) {
        R.x0 = r.x0; R.z0 = r.z0; R.nx = r.nx; R.nz = r.nz; R.ny = r.ny;
        for (int j = 0; j < r.ny; j++) R.y[j] = r.y[j];
        R.out_off = r.out_off;
        add_tiles(L.gt, (int)L.g.size(), r.nx, r.nz, tile_w, tile_h);
        L.g.push_back(R);
        L.max_ny_g = std::max(L.max_ny_g, (int)r.ny);
        L.columns += (int64_t)r.nx * r.nz;
        L.samples += (int64_t)r.nx * r.nz * r.ny;
        return;
    }
    // the zoom of block (x,y,z) picks one of the lattice corners ((x-2)>>2 + {0,1}, ...): the quart
    // cells of the region's blocks, at the two quart layers around each block y
    const int qx0 = (r.x0 - 2) >> 2, qx1 = ((r.x0 + r.nx - 1 - 2) >> 2) + 1;
    const int qz0 = (r.z0 - 2) >> 2, qz1 = ((r.z0 + r.nz - 1 - 2) >> 2) + 1;
    McbRequest Q = R;
    Q.x0 = qx0; Q.z0 = qz0; Q.nx = qx1 - qx0 + 1; Q.nz = qz1 - qz0 + 1; Q.ny = 2 * r.ny;
    for (int j = 0; j < r.ny; j++) {
        Q.y[2 * j] = (r.y[j] - 2) >> 2;
        Q.y[2 * j + 1] = ((r.y[j] - 2) >> 2) + 1;
    }
    Q.out_off = L.quart_bytes;
    L.quart_bytes += (int64_t)Q.nx * Q.nz * Q.ny;
    McbRequest Z = R;
    Z.x0 = r.x0; Z.z0 = r.z0; Z.nx = r.nx; Z.nz = r.nz; Z.ny = r.ny;
    for (int j = 0; j < r.ny; j++) Z.y[j] = r.y[j];
    Z.zoom = 1;
    Z.out_off = r.out_off;
    Z.hashed_seed = hashed_seed(r.seed);
    add_tiles(L.qt, (int)L.q.size(), Q.nx, Q.nz, tile_w, tile_h);
    L.max_ny_q = std::max(L.max_ny_q, (int)Q.ny);
    add_tiles(L.zt, (int)L.z.size(), Z.nx, Z.nz, 16, 16);
    L.q.push_back(Q);
    L.z.push_back(Z);
    L.columns += (int64_t)Q.nx * Q.nz;
    L.samples += (int64_t)r.nx * r.nz * r.ny;
}

// BiomeManager.obfuscateSeed: SHA-256 (FIPS 180-4) over the seed's 8 little-endian bytes, the first 8
// digest bytes read little-endian (libmcgen/src/world.c:118-152).  Host side, once per request.
int64_t hashed_seed(int64_t seed) {
    static const uint32_t K[64] = {
        0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5, 0xd807aa98,
        0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174, 0xe49b69c1, 0xefbe4786,
        0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da, 0x983e5152, 0xa831c66d, 0xb00327c8,
        0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967, 0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13,
        0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85, 0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819,
        0xd6990624, 0xf40e3585, 0x106aa070, 0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a,
        0x5b9cca4f, 0x682e6ff3, 0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7,
        0xc67178f2};
    auto rr = [](uint32_t v, int k) { return (v >> k) | (v << (32 - k)); };
    uint32_t w[64] = {0};
    const uint64_t u = (uint64_t)seed;
    uint8_t b[8];
    for (int i = 0; i < 8; i++) b[i] = (uint8_t)(u >> (8 * i));
    w[0] = (uint32_t)b[0] << 24 | (uint32_t)b[1] << 16 | (uint32_t)b[2] << 8 | b[3];
    w[1] = (uint32_t)b[4] << 24 | (uint32_t)b[5] << 16 | (uint32_t)b[6] << 8 | b[7];
    w[2] = 0x80000000u;
    w[15] = 64;
    for (int i = 16; i < 64; i++) {
        uint32_t s0 = rr(w[i - 15], 7) ^ rr(w[i - 15], 18) ^ (w[i - 15] >> 3);
        uint32_t s1 = rr(w[i - 2], 17) ^ rr(w[i - 2], 19) ^ (w[i - 2] >> 10);
        w[i] = w[i - 16] + s0 + w[i - 7] + s1;
    }
    const uint32_t H[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a, 0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    uint32_t a = H[0], bb = H[1], c = H[2], d = H[3], e = H[4], f = H[5], g = H[6], h = H[7];
    for (int i = 0; i < 64; i++) {
        uint32_t t1 = h + (rr(e, 6) ^ rr(e, 11) ^ rr(e, 25)) + ((e & f) ^ (~e & g)) + K[i] + w[i];
        uint32_t t2 = (rr(a, 2) ^ rr(a, 13) ^ rr(a, 22)) + ((a & bb) ^ (a & c) ^ (bb & c));
        h = g; g = f; f = e; e = d + t1; d = c; c = bb; bb = a; a = t1 + t2;
    }
    const uint32_t d0 = H[0] + a, d1 = H[1] + bb;
    const uint8_t dg[8] = {(uint8_t)(d0 >> 24), (uint8_t)(d0 >> 16), (uint8_t)(d0 >> 8), (uint8_t)d0,
                           (uint8_t)(d1 >> 24), (uint8_t)(d1 >> 16), (uint8_t)(d1 >> 8), (uint8_t)d1};
    uint64_t r = 0;
    for (int i = 7; i >= 0; i--) r = (r << 8) | dg[i];
    return (int64_t)r;
}

}  // namespace mcb

namespace {

// v into *d (grow-only: the buffer is kept while it holds v; an empty v leaves it as it is, unused)
template <class T>
int upload(T **d, size_t *cap, const std::vector<T> &v) {
    if (v.empty()) return 0;
    if (v.size() > *cap) {
        (void)hipFree(*d);
        *d = nullptr;
        *cap = 0;
        MCB_HIPCHK(hipMalloc((void **)d, v.size() * sizeof(T)));
        *cap = v.size();
    }
    MCB_HIPCHK(hipMemcpy(*d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

// One world's share of a plan.
struct Group {
    McbEngine *e = nullptr;
    std::vector<McbRegion> regs;        // for re-lowering by the autotuner
    std::vector<int> slot_of_reg;
    int nseeds = 0;
    int64_t *d_seeds = nullptr;
    int32_t *d_slot_of = nullptr;
    McbSlot *d_slots = nullptr;
    McbRequest *d_greqs = nullptr, *d_qreqs = nullptr, *d_zreqs = nullptr;
    McbTile *d_gtiles = nullptr, *d_qtiles = nullptr, *d_ztiles = nullptr;
    size_t cap_seeds = 0, cap_slot_of = 0, cap_slots = 0, cap_greqs = 0, cap_qreqs = 0, cap_zreqs = 0, cap_gtiles = 0,
           cap_qtiles = 0, cap_ztiles = 0;
    int ngtiles = 0, nqtiles = 0, nztiles = 0;
    uint8_t *d_quart = nullptr;
    int64_t quart_cap = 0;
    int tile_w = 16, cpl = 1, blocks_arg = 0, blocks_g = 0, blocks_q = 0;
    int max_ny_g = 1, max_ny_q = 1;
    bool nowrap_g = false, nowrap_q = false;       // wrap_free over the quart-grid / zoom-layer requests
    int64_t columns = 0, samples = 0;

    void release() {
        (void)hipFree(d_seeds); (void)hipFree(d_slot_of); (void)hipFree(d_slots);
        (void)hipFree(d_greqs); (void)hipFree(d_qreqs); (void)hipFree(d_zreqs);
        (void)hipFree(d_gtiles); (void)hipFree(d_qtiles); (void)hipFree(d_ztiles);
        (void)hipFree(d_quart);
        d_seeds = nullptr; d_slot_of = nullptr; d_slots = nullptr; d_greqs = d_qreqs = d_zreqs = nullptr;
        d_gtiles = d_qtiles = d_ztiles = nullptr; d_quart = nullptr;
        cap_seeds = cap_slot_of = cap_slots = cap_greqs = cap_qreqs = cap_zreqs = cap_gtiles = cap_qtiles = cap_ztiles = 0;
        quart_cap = 0;
    }

    // (re)build the descriptors for a tile shape
    int lower(int tw, int cp, int blocks) {
        Lowered L;
        for (size_t i = 0; i < regs.size(); i++) lower_region(regs[i], slot_of_reg[i], tw, cp, L);
        if ((int64_t)L.gt.size() > INT32_MAX || (int64_t)L.qt.size() > INT32_MAX) return fail(MCB_E_LIMIT, "too many tiles");
        int rc;
        if ((rc = upload(&d_greqs, &cap_greqs, L.g)) || (rc = upload(&d_gtiles, &cap_gtiles, L.gt)) ||
            (rc = upload(&d_qreqs, &cap_qreqs, L.q)) || (rc = upload(&d_qtiles, &cap_qtiles, L.qt)) ||
            (rc = upload(&d_zreqs, &cap_zreqs, L.z)) || (rc = upload(&d_ztiles, &cap_ztiles, L.zt)))
            return rc;
        if (L.quart_bytes > quart_cap) {
            (void)hipFree(d_quart);
            d_quart = nullptr;
            quart_cap = 0;
            if (hipMalloc((void **)&d_quart, (size_t)L.quart_bytes) != hipSuccess) return fail(MCB_E_NOMEM, "zoom staging");
            quart_cap = L.quart_bytes;
        }
        ngtiles = (int)L.gt.size();
        nqtiles = (int)L.qt.size();
        nztiles = (int)L.zt.size();
        tile_w = tw;
        cpl = cp;
        max_ny_g = L.max_ny_g;
        max_ny_q = L.max_ny_q;
        nowrap_g = wrap_free(e->w.host, L.g);
        nowrap_q = wrap_free(e->w.host, L.q);
        blocks_arg = blocks;
        blocks_g = blocks > 0 ? std::min(blocks, std::max(ngtiles, 1)) : ngtiles;
        blocks_q = blocks > 0 ? std::min(blocks, std::max(nqtiles, 1)) : nqtiles;
        columns = L.columns;
        samples = L.samples;
        return 0;
    }

    // lower() for a one-off call: the seeds, slot map and every descriptor list packed into one host image
    // and one copy into *d_pack (grow-only, *pack_cap bytes); the group's pointers point into it
    int lower_packed(const std::vector<int64_t> &seeds, const std::vector<int32_t> &ident, int tw, int cp, char **d_pack,
                     size_t *pack_cap) {
        Lowered L;
        for (size_t i = 0; i < regs.size(); i++) lower_region(regs[i], slot_of_reg[i], tw, cp, L);
        if ((int64_t)L.gt.size() > INT32_MAX || (int64_t)L.qt.size() > INT32_MAX) return fail(MCB_E_LIMIT, "too many tiles");
        struct Part { const void *src; size_t bytes; size_t off; };
        Part parts[] = {{seeds.data(), seeds.size() * sizeof(int64_t), 0}, {ident.data(), ident.size() * sizeof(int32_t), 0},
                        {L.g.data(), L.g.size() * sizeof(McbRequest), 0},  {L.gt.data(), L.gt.size() * sizeof(McbTile), 0},
                        {L.q.data(), L.q.size() * sizeof(McbRequest), 0},  {L.qt.data(), L.qt.size() * sizeof(McbTile), 0},
                        {L.z.data(), L.z.size() * sizeof(McbRequest), 0},  {L.zt.data(), L.zt.size() * sizeof(McbTile), 0}};
        size_t total = 0;
        for (auto &q : parts) { q.off = total; total += (q.bytes + 255) & ~(size_t)255; }
        total = std::max<size_t>(total, 256);
        if (total > *pack_cap) {
            (void)hipFree(*d_pack);
            *d_pack = nullptr;
            *pack_cap = 0;
            MCB_HIPCHK(hipMalloc((void **)d_pack, total));
            *pack_cap = total;
        }
        std::vector<char> image(total);
        for (auto &q : parts)
            if (q.bytes) memcpy(image.data() + q.off, q.src, q.bytes);
        MCB_HIPCHK(hipMemcpy(*d_pack, image.data(), total, hipMemcpyHostToDevice));
        d_seeds = (int64_t *)(*d_pack + parts[0].off);
        d_slot_of = (int32_t *)(*d_pack + parts[1].off);
        d_greqs = (McbRequest *)(*d_pack + parts[2].off);
        d_gtiles = (McbTile *)(*d_pack + parts[3].off);
        d_qreqs = (McbRequest *)(*d_pack + parts[4].off);
        d_qtiles = (McbTile *)(*d_pack + parts[5].off);
        d_zreqs = (McbRequest *)(*d_pack + parts[6].off);
        d_ztiles = (McbTile *)(*d_pack + parts[7].off);
        if (L.quart_bytes > quart_cap) {
            (void)hipFree(d_quart);
            d_quart = nullptr;
            quart_cap = 0;
            if (hipMalloc((void **)&d_quart, (size_t)L.quart_bytes) != hipSuccess) return fail(MCB_E_NOMEM, "zoom staging");
            quart_cap = L.quart_bytes;
        }
        ngtiles = (int)L.gt.size();
        nqtiles = (int)L.qt.size();
        nztiles = (int)L.zt.size();
        tile_w = tw;
        cpl = cp;
        max_ny_g = L.max_ny_g;
        max_ny_q = L.max_ny_q;
        nowrap_g = wrap_free(e->w.host, L.g);
        nowrap_q = wrap_free(e->w.host, L.q);
        blocks_arg = 0;
        blocks_g = ngtiles;
        blocks_q = nqtiles;
        columns = L.columns;
        samples = L.samples;
        return 0;
    }

    // limit > 0 restricts the grid stages to a prefix of the tile lists (autotuner trials)
    int stage(int s, uint8_t *d_out, hipStream_t st, int limit = 0) const {
        const DevWorld &w = e->w;
        const int ng = limit > 0 ? std::min(ngtiles, limit) : ngtiles, nq = limit > 0 ? std::min(nqtiles, limit) : nqtiles;
        const int bg = std::min(blocks_g, std::max(ng, 1)), bq = std::min(blocks_q, std::max(nq, 1));
        switch (s) {
            case 0: MCB_HIPCHK(launch_seed_init(w, d_seeds, d_slot_of, nseeds, d_slots, st)); break;
            case 1: MCB_HIPCHK(launch_grid(w, d_slots, d_greqs, d_gtiles, ng, d_out, tile_w, cpl, max_ny_g, bg, nowrap_g, st)); break;
            case 2: MCB_HIPCHK(launch_grid(w, d_slots, d_qreqs, d_qtiles, nq, d_quart, tile_w, cpl, max_ny_q, bq, nowrap_q, st)); break;
            default: MCB_HIPCHK(launch_zoom(d_zreqs, d_qreqs, d_ztiles, nztiles, d_quart, d_out, st)); break;
        }
        return 0;
    }
};

}  // namespace

struct McbPlan {
    std::vector<Group> groups;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    uint8_t *graph_out = nullptr;
    void drop_graph() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
        graph_out = nullptr;
    }
};

struct McbContext {
    std::string data_dir;
    struct Entry { int version, dim; std::string preset; McbEngine *e; };
    std::vector<Entry> worlds;
    std::mutex mu;
};

static int build_plan(const std::vector<McbEngine *> &engine_of_world, const McbRegion *reqs, int n, int tile_w, int blocks,
                      McbPlan **out) {
    // tile_w encodes the columns per lane in its upper bits: tile_w = width | cpl << 16 (cpl 1, 2 or 4)
    int cpl = (tile_w >> 16) & 0xFF;
    tile_w &= 0xFFFF;
    if (cpl == 0) cpl = 1;
    if (cpl < 1 || cpl > 64) return fail(MCB_E_ARG, "columns per lane must be 1..64");
    if (!reqs || n <= 0 || !out) return fail(MCB_E_ARG, "bad arguments");
    *out = nullptr;
    if (tile_w == 0) tile_w = 16;
    if (tile_w < 1 || tile_w > 256 || 256 % tile_w) return fail(MCB_E_ARG, "tile_w must divide 256");
    std::map<int, int> group_of_world;
    std::unique_ptr<McbPlan> p(new McbPlan());
    std::vector<std::unordered_map<int64_t, int>> seed_slot;
    std::vector<std::vector<int64_t>> seeds;
    for (int i = 0; i < n; i++) {
        const McbRegion &r = reqs[i];
        if (int rc = check_region(r)) return rc;
        const int wi = engine_of_world.size() == 1 ? 0 : r.world;
        if (wi < 0 || wi >= (int)engine_of_world.size() || !engine_of_world[wi]) return fail(MCB_E_ARG, "bad world index %d", r.world);
        auto it = group_of_world.find(wi);
        int gi;
        if (it == group_of_world.end()) {
            gi = (int)p->groups.size();
            group_of_world[wi] = gi;
            p->groups.emplace_back();
            p->groups.back().e = engine_of_world[wi];
            seed_slot.emplace_back();
            seeds.emplace_back();
        } else gi = it->second;
        Group &G = p->groups[gi];
        auto s = seed_slot[gi].find(r.seed);
        int slot;
        if (s == seed_slot[gi].end()) {
            slot = (int)seeds[gi].size();
            seed_slot[gi][r.seed] = slot;
            seeds[gi].push_back(r.seed);
        } else slot = s->second;
        G.regs.push_back(r);
        G.slot_of_reg.push_back(slot);
    }
    for (size_t gi = 0; gi < p->groups.size(); gi++) {
        Group &G = p->groups[gi];
        G.nseeds = (int)seeds[gi].size();
        std::vector<int32_t> ident(G.nseeds);
        for (int i = 0; i < G.nseeds; i++) ident[i] = i;
        int rc;
        if ((rc = upload(&G.d_seeds, &G.cap_seeds, seeds[gi])) || (rc = upload(&G.d_slot_of, &G.cap_slot_of, ident)) ||
            (rc = G.lower(tile_w, cpl, blocks))) {
            for (auto &g : p->groups) g.release();
            return rc;
        }
        if (hipMalloc((void **)&G.d_slots, (size_t)G.nseeds * sizeof(McbSlot)) != hipSuccess) {
            for (auto &g : p->groups) g.release();
            return fail(MCB_E_NOMEM, "state slots for %d seeds", G.nseeds);
        }
        G.cap_slots = (size_t)G.nseeds;
    }
    *out = p.release();
    return MCB_OK;
}

extern "C" {

int mcb_plan_create(McbEngine *e, const McbRegion *reqs, int n, int tile_w, int blocks, McbPlan **out) {
    if (!e) return fail(MCB_E_ARG, "null engine");
    return build_plan(std::vector<McbEngine *>{e}, reqs, n, tile_w, blocks, out);
}

int mcb_plan_run_stage(McbPlan *p, int stage, uint8_t *d_out, void *stream) {
    if (!p || !d_out || stage < 0 || stage > 3) return fail(MCB_E_ARG, "bad arguments");
    for (auto &G : p->groups)
        if (int rc = G.stage(stage, d_out, (hipStream_t)stream)) return rc;
    return MCB_OK;
}

int mcb_plan_run(McbPlan *p, uint8_t *d_out, void *stream) {
    if (!p || !d_out) return fail(MCB_E_ARG, "bad arguments");
    for (auto &G : p->groups)
        for (int s = 0; s < 4; s++)
            if (int rc = G.stage(s, d_out, (hipStream_t)stream)) return rc;
    return MCB_OK;
}

int mcb_plan_run_graph(McbPlan *p, uint8_t *d_out, void *stream) {
    if (!p || !d_out) return fail(MCB_E_ARG, "bad arguments");
    if (p->gexec && p->graph_out != d_out) p->drop_graph();
    if (!p->gexec) {
        hipStream_t cs;
        MCB_HIPCHK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        hipError_t be = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        if (be != hipSuccess) {
            (void)hipStreamDestroy(cs);
            return fail(MCB_E_DEVICE, "begin capture: %s", hipGetErrorString(be));
        }
        int rc = mcb_plan_run(p, d_out, cs);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(cs, &g);
        (void)hipStreamDestroy(cs);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ce != hipSuccess) return fail(MCB_E_DEVICE, "graph capture: %s", hipGetErrorString(ce));
        p->graph = g;
        hipError_t ie = hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
            p->drop_graph();
            return fail(MCB_E_DEVICE, "graph instantiate: %s", hipGetErrorString(ie));
        }
        p->graph_out = d_out;
    }
    MCB_HIPCHK(hipGraphLaunch(p->gexec, (hipStream_t)stream));
    return MCB_OK;
}

int mcb_plan_info(const McbPlan *p, int64_t *tiles, int64_t *columns, int64_t *samples, int64_t *seeds) {
    if (!p) return fail(MCB_E_ARG, "null plan");
    int64_t t = 0, c = 0, s = 0, n = 0;
    for (auto &G : p->groups) { t += G.ngtiles + G.nqtiles; c += G.columns; s += G.samples; n += G.nseeds; }
    if (tiles) *tiles = t;
    if (columns) *columns = c;
    if (samples) *samples = s;
    if (seeds) *seeds = n;
    return MCB_OK;
}

int mcb_plan_shape(const McbPlan *p, int group, int *tile_w, int *cpl, int *blocks, int *wrap_free_grid) {
    if (!p || group < 0 || group >= (int)p->groups.size()) return fail(MCB_E_ARG, "no plan group %d", group);
    const Group &G = p->groups[group];
    if (tile_w) *tile_w = G.tile_w;
    if (cpl) *cpl = G.cpl;
    if (blocks) *blocks = G.blocks_arg;
    // the grid launch takes the wrap-free kernel: 26.3, or 26.1/26.2 up to two layers (launch_grid)
    if (wrap_free_grid) *wrap_free_grid = G.nowrap_g && (G.e->w.host.fmt_new || G.max_ny_g <= 2) ? 1 : 0;
    return MCB_OK;
}

void mcb_plan_free(McbPlan *p) {
    if (!p) return;
    p->drop_graph();
    for (auto &G : p->groups) G.release();
    delete p;
}

// trial launches run a prefix of each tile list, long enough to fill the device many times over; the
// whole-plan time is estimated as trial time x (tiles / trial tiles)
// Autotuner trials: a prefix of each grid launch holding about this many columns, the same work for every
// tile shape (tiles hold 256 x cols_per_lane columns), timed and scaled to the whole launch.
static const int64_t kTuneColumns = 1 << 24;

int mcb_plan_autotune(McbPlan *p, uint8_t *d_out, void *stream, int *tile_w_out, double *best_ms_out) {
    if (!p || !d_out) return fail(MCB_E_ARG, "bad arguments");
    hipStream_t st = (hipStream_t)stream;
    p->drop_graph();
    int cus = 256;
    {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
    }
    hipEvent_t e0, e1;
    MCB_HIPCHK(hipEventCreate(&e0));
    MCB_HIPCHK(hipEventCreate(&e1));
    double total_best = 0.0;
    for (auto &G : p->groups) {
        if (int rc = G.stage(0, d_out, st)) return rc;
        const int shapes[4] = {16, 32, 64, 8};
        const int cpls[4] = {1, 2, 4, 16};
        const int per_cu[3] = {0, 2, 4};      // 0 = one workgroup per tile
        double best = 1e30;
        int best_tw = G.tile_w, best_cpl = G.cpl, best_blocks = G.blocks_arg;
        for (int tw : shapes) {
            for (int cp : cpls) {
                for (int k : per_cu) {
                    const int blocks = k ? k * cus : 0;
                    if (int rc = G.lower(tw, cp, blocks)) return rc;
                    const int trial = (int)std::max<int64_t>(1, kTuneColumns / (256 * cp));
                    float ms = 0.f, acc = 0.f;
                    for (int rep = 0; rep < 3; rep++) {    // first rep warms caches, keep the best of the rest
                        MCB_HIPCHK(hipEventRecord(e0, st));
                        if (int rc = G.stage(1, d_out, st, trial)) return rc;
                        if (int rc = G.stage(2, d_out, st, trial)) return rc;
                        MCB_HIPCHK(hipEventRecord(e1, st));
                        MCB_HIPCHK(hipEventSynchronize(e1));
                        MCB_HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                        if (rep == 1 || (rep > 1 && ms < acc)) acc = ms;
                    }
                    acc *= (float)(G.ngtiles + G.nqtiles) / (float)std::max(1, std::min(G.ngtiles, trial) + std::min(G.nqtiles, trial));
                    if (acc < best) { best = acc; best_tw = tw; best_cpl = cp; best_blocks = blocks; }
                }
            }
        }
        if (int rc = G.lower(best_tw, best_cpl, best_blocks)) return rc;
        total_best += best;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (tile_w_out) *tile_w_out = p->groups.empty() ? 0 : (p->groups[0].tile_w | p->groups[0].cpl << 16);
    if (best_ms_out) *best_ms_out = total_best;
    return MCB_OK;
}

// One-off regions: the group a one-region plan would hold, kept per engine and rebuilt for each call with
// its descriptors in one grow-only buffer (one upload) and a grow-only output buffer (a plan made and
// freed per call, with an allocation and a copy per descriptor list, took 0.162 ms for a 256 x 256
// region).  The same descriptors and launches as mcb_plan_create(engine, &region, 1, 16, 0) +
// mcb_plan_run.
struct RegionCache {
    Group G;                            // its descriptor pointers point into d_pack (not freed separately)
    char *d_pack = nullptr;
    size_t pack_cap = 0;
    uint8_t *d_out = nullptr;
    size_t out_cap = 0;
};
static void region_cache_free(void *p) {
    RegionCache *c = static_cast<RegionCache *>(p);
    (void)hipDeviceSynchronize();
    (void)hipFree(c->G.d_slots);
    (void)hipFree(c->G.d_quart);
    (void)hipFree(c->d_pack);
    (void)hipFree(c->d_out);
    delete c;
}

int mcb_region(McbEngine *e, const McbRegion *req, uint8_t *out_host) {
    if (!e || !req || !out_host) return fail(MCB_E_ARG, "bad arguments");
    McbRegion r = *req;
    r.out_off = 0;
    if (int rc = check_region(r)) return rc;
    std::lock_guard<std::mutex> lk(e->region_mu);
    if (!e->region_cache) {
        RegionCache *c = new RegionCache();
        c->G.e = e;
        e->region_cache = c;
        e->region_free = region_cache_free;
    }
    RegionCache &c = *static_cast<RegionCache *>(e->region_cache);
    Group &G = c.G;
    G.regs.assign(1, r);
    G.slot_of_reg.assign(1, 0);
    G.nseeds = 1;
    const std::vector<int64_t> seeds{r.seed};
    const std::vector<int32_t> ident{0};
    if (int rc = G.lower_packed(seeds, ident, 16, 1, &c.d_pack, &c.pack_cap)) return rc;
    if (G.cap_slots < 1) {
        if (hipMalloc((void **)&G.d_slots, sizeof(McbSlot)) != hipSuccess) return fail(MCB_E_NOMEM, "state slot");
        G.cap_slots = 1;
    }
    const size_t bytes = (size_t)r.ny * r.nx * r.nz;
    if (bytes > c.out_cap) {
        (void)hipFree(c.d_out);
        c.d_out = nullptr;
        c.out_cap = 0;
        if (hipMalloc((void **)&c.d_out, bytes) != hipSuccess) return fail(MCB_E_NOMEM, "output");
        c.out_cap = bytes;
    }
    for (int s = 0; s < 4; s++)
        if (int rc = G.stage(s, c.d_out, nullptr)) return rc;
    if (mcb::copy_to_host(out_host, c.d_out, bytes, nullptr) != hipSuccess) return fail(MCB_E_DEVICE, "copy back");
    return MCB_OK;
}

int64_t mcb_regions_bytes(const McbRegion *r, int n) {
    int64_t m = 0;
    for (int i = 0; i < n; i++) m = std::max(m, r[i].out_off + (int64_t)r[i].ny * r[i].nx * r[i].nz);
    return m;
}

int mcb_context_create(McbContext **out, const char *data_dir) {
    if (!out) return fail(MCB_E_ARG, "null out");
    McbContext *c = new McbContext();
    c->data_dir = data_dir ? data_dir : default_data_dir();
    *out = c;
    return MCB_OK;
}

void mcb_context_free(McbContext *c) {
    if (!c) return;
    for (auto &w : c->worlds) mcb_close(w.e);
    delete c;
}

int mcb_context_world(McbContext *c, int version, int dim, const char *preset) {
    if (!c) return -fail(MCB_E_ARG, "null context");
    std::string p = (!preset || !strcmp(preset, "normal") || dim != MCB_OVERWORLD) ? "overworld" : preset;
    std::lock_guard<std::mutex> lk(c->mu);
    for (size_t i = 0; i < c->worlds.size(); i++)
        if (c->worlds[i].version == version && c->worlds[i].dim == dim && c->worlds[i].preset == p) return (int)i;
    McbEngine *e = nullptr;
    if (int rc = mcb_open(&e, version, dim, p.c_str(), c->data_dir.c_str(), 0)) return -rc;
    c->worlds.push_back({version, dim, p, e});
    return (int)c->worlds.size() - 1;
}

int mcb_context_world_engine(McbContext *c, int world, McbEngine **e) {
    if (!c || !e) return fail(MCB_E_ARG, "bad arguments");
    std::lock_guard<std::mutex> lk(c->mu);
    if (world < 0 || world >= (int)c->worlds.size()) return fail(MCB_E_ARG, "bad world index %d", world);
    *e = c->worlds[world].e;
    return MCB_OK;
}

int mcb_context_plan(McbContext *c, const McbRegion *reqs, int n, int tile_w, int blocks, McbPlan **out) {
    if (!c) return fail(MCB_E_ARG, "null context");
    std::vector<McbEngine *> ew;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        for (auto &w : c->worlds) ew.push_back(w.e);
    }
    if (ew.empty()) return fail(MCB_E_ARG, "context has no worlds");
    if (ew.size() == 1) ew.push_back(nullptr);      // keep world indexing even with one world
    return build_plan(ew, reqs, n, tile_w, blocks, out);
}

}  // extern "C"
