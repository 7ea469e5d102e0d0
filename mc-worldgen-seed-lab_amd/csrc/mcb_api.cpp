// mcb_api.cpp - engine objects, the points path and the reference CPU-engine drop-in
// (include/mcbiome.h).  Host code only: data loading goes through build_world (mcb_world.cpp), all
// sampling through the HIP kernels (mcb_kernels.hip).  There is no CPU compute path: a missing device
// is an error.
#include <dlfcn.h>

#include <algorithm>
#include <map>
#include <memory>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mcb_internal.h"

using namespace mcb;

namespace mcb {
static thread_local std::string g_err;

int failv(int code, const char *fmt, va_list ap) {
    char buf[512];
    vsnprintf(buf, sizeof buf, fmt, ap);
    g_err = buf;
    return code;
}

int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    failv(code, fmt, ap);
    va_end(ap);
    return code;
}

std::string default_data_dir() {
    if (const char *d = getenv("MCB_DATA")) return d;
    if (const char *r = getenv("MCGEN_ROOT")) return std::string(r) + "/data";
    Dl_info info;
    if (dladdr((void *)&default_data_dir, &info) && info.dli_fname) {
        std::string p = info.dli_fname;
        size_t s = p.rfind('/');
        std::string dir = s == std::string::npos ? "." : p.substr(0, s);
        return dir + "/../data";      // <pkg>/lib/libmcbiome_hip.so -> <pkg>/data
    }
    return "data";
}
}  // namespace mcb

namespace {
int device_ready() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(MCB_E_DEVICE, "no HIP device");
    return 0;
}

}  // namespace

// identity slot map (slot i = seed i) of at least n entries
int mcb::ensure_ident(McbEngine *e, int n) {
    if ((size_t)n <= e->ident_cap) return 0;
    if (int rc = dev_grow(&e->d_ident, &e->ident_cap, (size_t)n)) return rc;
    std::vector<int32_t> id(n);
    for (int i = 0; i < n; i++) id[i] = i;
    MCB_HIPCHK(hipMemcpy(e->d_ident, id.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
    return 0;
}

// Cell tables depend only on the tree (not on the seed or preset): one build per distinct tree (and
// dominance rule) per process, shared by every engine, context world and plugin world that loads it.
// The key hashes the tree eight bytes at a time (McbNode is 32 bytes), FNV-1a style over 64-bit words.
// max_cells > 0: on a miss, a tree whose breakpoint grid would exceed that many cells gets no table
// (remembered as such).
std::shared_ptr<const CellTable> mcb::cells_for_tree(const std::vector<McbNode> &tree, std::string &err, bool strict,
                                                     double max_cells) {
    static std::mutex mu;
    static std::map<std::pair<uint64_t, uint64_t>, std::shared_ptr<const CellTable>> cache;
    static_assert(sizeof(McbNode) % 8 == 0, "McbNode words");
    uint64_t h = 1469598103934665603ULL ^ (strict ? 0x5bd1e995ULL : 0);
    const uint64_t *p = reinterpret_cast<const uint64_t *>(tree.data());
    for (size_t i = 0; i < tree.size() * sizeof(McbNode) / 8; i++) h = (h ^ p[i]) * 1099511628211ULL;
    const std::pair<uint64_t, uint64_t> key(h, (uint64_t)tree.size() << 1 | (strict ? 1 : 0));
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
        if (!it->second) err = "cell table over the size limit";
        return it->second;
    }
    if (max_cells > 0) {
        std::vector<int32_t> ends[6];
        for (const McbNode &n : tree)
            if (n.leaf & 0x8000)
                for (int k = 0; k < 6; k++) { ends[k].push_back(n.lo[k]); ends[k].push_back(n.hi[k]); }
        double cells = 1.0;
        for (int k = 0; k < 6; k++) {
            std::sort(ends[k].begin(), ends[k].end());
            cells *= (double)(std::unique(ends[k].begin(), ends[k].end()) - ends[k].begin()) + 1.0;
        }
        if (cells > max_cells) {
            cache[key] = nullptr;
            err = "cell table over the size limit";
            return nullptr;
        }
    }
    auto ct = std::make_shared<CellTable>();
    if (build_cells(tree, *ct, err, strict) != 0) return nullptr;
    cache[key] = ct;
    return ct;
}

namespace {

// Fills the world's cell-table fields and uploads the entries and lists.
int attach_cells(McbEngine *e, const std::vector<McbNode> &tree) {
    std::string err;
    auto ct = cells_for_tree(tree, err);
    if (!ct) return fail(MCB_E_DATA, "cell table: %s", err.c_str());
    McbWorld &W = e->w.host;
    int n[6];
    for (int k = 0; k < 6; k++) n[k] = ct->nb[k] + 1;
    // index = (((((cT*nH + cH)*nC + cC)*nE + cE)*nW + cW)*nD + cD)
    W.cell_stride[4] = 1;
    W.cell_stride[5] = n[4];
    W.cell_stride[3] = n[5] * W.cell_stride[5];
    W.cell_stride[2] = n[3] * W.cell_stride[3];
    W.cell_stride[1] = n[2] * W.cell_stride[2];
    W.cell_stride[0] = n[1] * W.cell_stride[1];
    if ((size_t)n[0] * W.cell_stride[0] != ct->entry.size()) return fail(MCB_E_DATA, "cell table size mismatch");
    // locator buckets of 256 values; the table is used only when no bucket holds two breakpoints and
    // every cell offset fits 20 bits
    W.cell_on = ct->entry.size() <= (1u << 20);
    for (int k = 0; k < 6; k++) {
        const int32_t *bp = ct->bp.data() + ct->off[k];
        for (int b = 0; b < 256; b++) {
            const int t0 = b * 256 - 32768;
            int below = 0, inside = 0, off = 256;
            for (int i = 0; i < ct->nb[k]; i++) {
                if (bp[i] <= t0) below++;
                else if (bp[i] <= t0 + 255) {
                    inside++;
                    off = bp[i] - t0;
                }
            }
            if (inside > 1 || (ct->nb[k] && (bp[0] <= -32768 || bp[ct->nb[k] - 1] > 32767))) W.cell_on = 0;
            W.cell_lut[k][b] = (uint32_t)below * (uint32_t)W.cell_stride[k] | (uint32_t)off << 20;
        }
    }
    e->w.cells_build_ms = ct->build_ms;
    // device form: the listed leaves copied in list order (one load per candidate, no index hop); an
    // entry whose leaves all carry one biome answers without any distance: (biome << 8) | 1
    std::vector<McbNode> recs(std::max<size_t>(ct->list.size(), 1));
    for (size_t i = 0; i < ct->list.size(); i++) recs[i] = tree[ct->list[i]];
    std::vector<uint32_t> dent(ct->entry.size());
    for (size_t c = 0; c < ct->entry.size(); c++) {
        const uint32_t en = ct->entry[c], cnt = en & 255u, first = en >> 8;
        bool uniform = cnt > 0;
        for (uint32_t q = 1; q < cnt && uniform; q++)
            uniform = (recs[first + q].leaf & 0xFF) == (recs[first].leaf & 0xFF);
        dent[c] = uniform ? (uint32_t)(recs[first].leaf & 0xFF) << 8 | 1u : en;
    }
    // the union over each five-dimension cell's depth cells (ascending pre-order, so a strict < still keeps
    // the first least-distance leaf): the layers of one column share those five dimensions, and one scan
    // of the union serves every layer.  A union that would exceed 255 leaves, or that meets a tree-walk
    // entry, gets count 0.
    const int nd = n[4];
    const size_t n5 = ct->entry.size() / (size_t)nd;
    std::vector<uint32_t> uent(n5);
    std::vector<McbNode> urecs;
    std::vector<uint16_t> merged;
    for (size_t c5 = 0; c5 < n5; c5++) {
        merged.clear();
        bool ok = true;
        for (int d = 0; d < nd && ok; d++) {
            const uint32_t en = ct->entry[c5 * nd + d], cnt = en & 255u, first = en >> 8;
            ok = cnt > 0;
            for (uint32_t q = 0; q < cnt; q++) merged.push_back(ct->list[first + q]);
        }
        std::sort(merged.begin(), merged.end());
        merged.erase(std::unique(merged.begin(), merged.end()), merged.end());
        if (!ok || merged.size() > 255 || urecs.size() >= (1u << 24)) {
            uent[c5] = 0;
            continue;
        }
        uent[c5] = (uint32_t)urecs.size() << 8 | (uint32_t)merged.size();
        for (uint16_t l : merged) urecs.push_back(tree[l]);
    }
    if (urecs.empty()) urecs.resize(1);
    if (hipMalloc((void **)&e->w.d_cells, dent.size() * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc((void **)&e->w.d_cell_recs, recs.size() * sizeof(McbNode)) != hipSuccess ||
        hipMalloc((void **)&e->w.d_cell_u, uent.size() * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc((void **)&e->w.d_cell_urecs, urecs.size() * sizeof(McbNode)) != hipSuccess)
        return fail(MCB_E_NOMEM, "cell table allocation");
    MCB_HIPCHK(hipMemcpy(e->w.d_cells, dent.data(), dent.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    MCB_HIPCHK(hipMemcpy(e->w.d_cell_recs, recs.data(), recs.size() * sizeof(McbNode), hipMemcpyHostToDevice));
    MCB_HIPCHK(hipMemcpy(e->w.d_cell_u, uent.data(), uent.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    MCB_HIPCHK(hipMemcpy(e->w.d_cell_urecs, urecs.data(), urecs.size() * sizeof(McbNode), hipMemcpyHostToDevice));
    e->cell_entries = ct->entry.size();
    e->cell_list_len = ct->list.size();
    e->cells = ct;
    return MCB_OK;
}

// seeds are processed in chunks so the points grid stays within the y-dimension limit
const int kSeedChunk = 32768;

int points_locked(McbEngine *e, const int64_t *d_seeds, int ns, const int32_t *d_pts, int np, uint8_t *d_out,
                  int32_t *d_tg, hipStream_t st) {
    const int chunk = std::min(ns, kSeedChunk);
    if (int rc = dev_grow(&e->d_slots, &e->slots_cap, (size_t)chunk)) return rc;
    if (int rc = ensure_ident(e, chunk)) return rc;
    for (int s0 = 0; s0 < ns; s0 += chunk) {
        const int n = std::min(chunk, ns - s0);
        const size_t o = (size_t)s0 * np;
        MCB_HIPCHK(launch_seed_init(e->w, d_seeds + s0, e->d_ident, n, e->d_slots, st));
        MCB_HIPCHK(launch_points(e->w, e->d_slots, e->d_ident, n, d_pts, np, d_out + o, d_tg ? d_tg + o * 6 : nullptr, st));
    }
    return MCB_OK;
}
}  // namespace

extern "C" {

const char *mcb_last_error(void) { return g_err.c_str(); }

int mcb_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int mcg_version_from_name(const char *s) {
    if (!s) return -1;
    if (!strncmp(s, "26.1", 4)) return MCB_V26_1;
    if (!strncmp(s, "26.2", 4)) return MCB_V26_2;
    if (!strncmp(s, "26.3", 4)) return MCB_V26_3;
    if (!strncmp(s, "26.4", 4)) return MCB_V26_4;
    return -1;
}
const char *mcg_version_name(int v) {
    return v == 0 ? "26.1" : v == 1 ? "26.2" : v == 2 ? "26.3" : v == 3 ? "26.4-snapshot-2" : "?";
}
int mcg_biome_count(void) { return biome_count(); }
const char *mcg_biome_name(int id) { return biome_name(id); }
int mcg_biome_id(const char *name) { return name ? biome_id(name) : -1; }

int mcb_tree(int version, int dim, const char *data_dir, int32_t *out, int cap) {
    if (dim != MCB_OVERWORLD && dim != MCB_NETHER) return -fail(MCB_E_ARG, "the End has no biome tree");
    McbWorld w;
    std::vector<McbNode> tree;
    std::string err;
    if (build_world(data_dir ? data_dir : default_data_dir(), version, dim, "normal", &w, tree, err) != 0)
        return -fail(MCB_E_DATA, "%s", err.c_str());
    if (out) {
        if ((int64_t)tree.size() * 16 > cap) return -fail(MCB_E_LIMIT, "%zu nodes need %zu ints", tree.size(), tree.size() * 16);
        for (size_t i = 0; i < tree.size(); i++) {
            int32_t *o = out + 16 * i;
            const McbNode &n = tree[i];
            for (int d = 0; d < 6; d++) { o[2 * d] = n.lo[d]; o[2 * d + 1] = n.hi[d]; }
            o[12] = (int32_t)n.d6;
            o[13] = n.escape;
            o[14] = n.leaf ? (n.leaf & 0xFF) : -1;
            o[15] = 0;
        }
    }
    return (int)tree.size();
}

int mcb_open(McbEngine **out, int version, int dim, const char *preset, const char *data_dir, int max_slots) {
    (void)max_slots;
    if (!out) return fail(MCB_E_ARG, "null out");
    *out = nullptr;
    if (int rc = device_ready()) return rc;
    McbEngine *e = new McbEngine();
    std::string err;
    std::vector<McbNode> tree;
    std::string root = data_dir ? data_dir : default_data_dir();
    if (build_world(root, version, dim, preset ? preset : "normal", &e->w.host, tree, err) != 0) {
        delete e;
        return fail(MCB_E_DATA, "%s", err.c_str());
    }
    e->w.version = version;
    e->w.dim = dim;
    e->w.preset = preset ? preset : "normal";
    // the cell table encodes closed parameter intervals; 26.4-snapshot-2's exclusive upper bounds (boxes
    // with lo = hi + 1) are searched by the tree walk alone
    if (dim == MCB_OVERWORLD && version < MCB_V26_4) {
        if (int rc = attach_cells(e, tree)) {
            mcb_close(e);
            return rc;
        }
    }
    if (hipGetDevice(&e->device) != hipSuccess || hipMalloc((void **)&e->w.d_world, sizeof(McbWorld)) != hipSuccess ||
        hipMemcpy(e->w.d_world, &e->w.host, sizeof(McbWorld), hipMemcpyHostToDevice) != hipSuccess) {
        mcb_close(e);
        return fail(MCB_E_DEVICE, "world upload failed");
    }
    e->tree_host = tree;
    if (!tree.empty()) {
        if (hipMalloc((void **)&e->w.d_tree, tree.size() * sizeof(McbNode)) != hipSuccess ||
            hipMemcpy(e->w.d_tree, tree.data(), tree.size() * sizeof(McbNode), hipMemcpyHostToDevice) != hipSuccess) {
            mcb_close(e);
            return fail(MCB_E_DEVICE, "tree upload failed");
        }
    }
    *out = e;
    return MCB_OK;
}

void mcb_close(McbEngine *e) {
    if (!e) return;
    if (e->region_cache && e->region_free) e->region_free(e->region_cache);
    (void)hipFree(e->w.d_world);
    (void)hipFree(e->w.d_tree);
    (void)hipFree(e->w.d_cells);
    (void)hipFree(e->w.d_cell_recs);
    (void)hipFree(e->w.d_cell_u);
    (void)hipFree(e->w.d_cell_urecs);
    (void)hipFree(e->d_slots);
    (void)hipFree(e->d_ident);
    (void)hipFree(e->d_seeds);
    (void)hipFree(e->d_pts);
    (void)hipFree(e->d_out);
    (void)hipFree(e->d_tg);
    (void)hipFree(e->d_field);
    (void)hipFree(e->d_masks);
    (void)hipFree(e->d_raw);
    (void)hipFree(e->d_xyzd);
    (void)hipFree(e->d_feas);
    (void)hipFree(e->d_coord_lut);
    (void)hipFree(e->d_obs);
    (void)hipFree(e->d_hist);
    (void)hipFree(e->d_next);
    (void)hipFree(e->d_nfound);
    (void)hipFree(e->d_found);
    (void)hipFree(e->d_list);
    delete e;
}

int mcb_engine_info(const McbEngine *e, int *version, int *dim, int *mode, int *tree_nodes) {
    if (!e) return fail(MCB_E_ARG, "null engine");
    if (version) *version = e->w.version;
    if (dim) *dim = e->w.dim;
    if (mode) *mode = e->w.host.fmt_new;
    if (tree_nodes) *tree_nodes = e->w.host.tree_n;
    return MCB_OK;
}

int mcb_engine_cells(const McbEngine *e, int64_t *cells, int64_t *list_len, double *build_ms) {
    if (!e) return fail(MCB_E_ARG, "null engine");
    if (cells) *cells = (int64_t)e->cell_entries;
    if (list_len) *list_len = (int64_t)e->cell_list_len;
    if (build_ms) *build_ms = e->w.cells_build_ms;
    return MCB_OK;
}

int mcb_targets(McbEngine *e, const int32_t *tg6, int n, uint8_t *out_biome) {
    if (!e || n < 0 || (n && (!tg6 || !out_biome))) return fail(MCB_E_ARG, "bad arguments");
    if (e->w.dim == MCB_END) return fail(MCB_E_ARG, "the End has no climate targets");
    for (int64_t i = 0; i < (int64_t)n * 6; i++)
        if (tg6[i] < -MCB_TARGET_MAX || tg6[i] > MCB_TARGET_MAX) return fail(MCB_E_ARG, "target %lld out of range", (long long)i);
    if (n == 0) return MCB_OK;
    std::lock_guard<std::mutex> lk(e->mu);
    if (int rc = dev_grow(&e->d_tg, &e->tg_cap, (size_t)n * 6)) return rc;
    if (int rc = dev_grow(&e->d_out, &e->out_cap, (size_t)n)) return rc;
    MCB_HIPCHK(hipMemcpy(e->d_tg, tg6, (size_t)n * 6 * sizeof(int32_t), hipMemcpyHostToDevice));
    MCB_HIPCHK(launch_targets(e->w, e->d_tg, n, e->d_out, nullptr));
    MCB_HIPCHK(hipMemcpy(out_biome, e->d_out, (size_t)n, hipMemcpyDeviceToHost));
    return MCB_OK;
}

int mcb_tie_masks(McbEngine *e, const int32_t *tg6, int n, uint64_t *out_masks) {
    if (!e || n < 0 || (n && (!tg6 || !out_masks))) return fail(MCB_E_ARG, "bad arguments");
    if (e->w.dim == MCB_END) return fail(MCB_E_ARG, "the End has no climate targets");
    if (n == 0) return MCB_OK;
    std::lock_guard<std::mutex> lk(e->mu);
    if (int rc = dev_grow(&e->d_tg, &e->tg_cap, (size_t)n * 6)) return rc;
    if (int rc = dev_grow(&e->d_masks, &e->masks_cap, (size_t)n * 2)) return rc;
    MCB_HIPCHK(hipMemcpy(e->d_tg, tg6, (size_t)n * 6 * sizeof(int32_t), hipMemcpyHostToDevice));
    MCB_HIPCHK(launch_tie_masks(e->w, e->d_tg, n, e->d_masks, nullptr));
    MCB_HIPCHK(hipMemcpy(out_masks, e->d_masks, (size_t)n * 2 * sizeof(uint64_t), hipMemcpyDeviceToHost));
    return MCB_OK;
}

}  // extern "C"

namespace {
// Seeds in chunks of at most kSeedChunk through the engine's slots: init, then `body(first seed, count)`.
template <class F>
int per_seed_chunks(McbEngine *e, const int64_t *seeds, int ns, F body) {
    const int chunk = std::min(ns, kSeedChunk);
    if (int rc = dev_grow(&e->d_slots, &e->slots_cap, (size_t)chunk)) return rc;
    if (int rc = ensure_ident(e, chunk)) return rc;
    if (int rc = dev_grow(&e->d_seeds, &e->seeds_cap, (size_t)chunk)) return rc;
    for (int s0 = 0; s0 < ns; s0 += chunk) {
        const int n = std::min(chunk, ns - s0);
        MCB_HIPCHK(hipMemcpy(e->d_seeds, seeds + s0, (size_t)n * sizeof(int64_t), hipMemcpyHostToDevice));
        MCB_HIPCHK(launch_seed_init(e->w, e->d_seeds, e->d_ident, n, e->d_slots, nullptr));
        if (int rc = body(s0, n)) return rc;
    }
    return MCB_OK;
}
}  // namespace

extern "C" {

int mcb_climate_raw(McbEngine *e, const int64_t *seeds, int ns, const int32_t *xyz, int np, float *out6) {
    if (!e || ns < 0 || np < 0 || (ns && np && (!seeds || !xyz || !out6))) return fail(MCB_E_ARG, "bad arguments");
    if (e->w.dim == MCB_END) return fail(MCB_E_ARG, "the End has no climate noises");
    if (ns == 0 || np == 0) return MCB_OK;
    std::lock_guard<std::mutex> lk(e->mu);
    const int chunk = std::min(ns, kSeedChunk);
    if (int rc = dev_grow(&e->d_pts, &e->pts_cap, (size_t)np * 3)) return rc;
    if (int rc = dev_grow(&e->d_raw, &e->raw_cap, (size_t)chunk * np * 6)) return rc;
    MCB_HIPCHK(hipMemcpy(e->d_pts, xyz, (size_t)np * 3 * sizeof(int32_t), hipMemcpyHostToDevice));
    return per_seed_chunks(e, seeds, ns, [&](int s0, int n) -> int {
        MCB_HIPCHK(launch_climate_raw(e->w, e->d_slots, n, e->d_pts, np, e->d_raw, nullptr));
        MCB_HIPCHK(hipMemcpy(out6 + (size_t)s0 * np * 6, e->d_raw, (size_t)n * np * 6 * sizeof(float), hipMemcpyDeviceToHost));
        return MCB_OK;
    });
}

int mcb_noise_values(McbEngine *e, int noise, const int64_t *seeds, int ns, const double *xyz, int np, double *out) {
    if (!e || ns < 0 || np < 0 || (ns && np && (!seeds || !xyz || !out))) return fail(MCB_E_ARG, "bad arguments");
    if (e->w.dim == MCB_END) return fail(MCB_E_ARG, "the End has no climate noises");
    if (noise < 0 || noise >= e->w.host.n_noise) return fail(MCB_E_ARG, "noise %d: this world has %d", noise, e->w.host.n_noise);
    if (ns == 0 || np == 0) return MCB_OK;
    std::lock_guard<std::mutex> lk(e->mu);
    const int chunk = std::min(ns, kSeedChunk);
    if (int rc = dev_grow(&e->d_xyzd, &e->xyzd_cap, (size_t)np * 3)) return rc;
    if (int rc = dev_grow(&e->d_field, &e->field_cap, (size_t)chunk * np)) return rc;
    MCB_HIPCHK(hipMemcpy(e->d_xyzd, xyz, (size_t)np * 3 * sizeof(double), hipMemcpyHostToDevice));
    return per_seed_chunks(e, seeds, ns, [&](int s0, int n) -> int {
        MCB_HIPCHK(launch_noise_values(e->w, e->d_slots, n, noise, e->d_xyzd, np, e->d_field, nullptr));
        MCB_HIPCHK(hipMemcpy(out + (size_t)s0 * np, e->d_field, (size_t)n * np * sizeof(double), hipMemcpyDeviceToHost));
        return MCB_OK;
    });
}

int mcb_end_field(McbEngine *e, const int64_t *seeds, int ns, int bx0, int bz0, int nx, int nz, int step, int kind,
                  double *out) {
    if (!e || ns < 0 || nx < 0 || nz < 0 || step < 1 || (kind != 0 && kind != 1) || (ns && nx && nz && (!seeds || !out)))
        return fail(MCB_E_ARG, "bad arguments");
    if (e->w.dim != MCB_END) return fail(MCB_E_ARG, "the End field needs an End engine");
    if (ns == 0 || nx == 0 || nz == 0) return MCB_OK;
    const int64_t xl = (int64_t)bx0 + (int64_t)(nx - 1) * step, zl = (int64_t)bz0 + (int64_t)(nz - 1) * step;
    if (xl > INT32_MAX || zl > INT32_MAX) return fail(MCB_E_LIMIT, "grid beyond the int32 block range");
    if ((int64_t)nx * nz > (1LL << 30)) return fail(MCB_E_LIMIT, "grid too large");
    std::lock_guard<std::mutex> lk(e->mu);
    const int chunk = (int)std::min<int64_t>(ns, std::max<int64_t>(1, std::min<int64_t>(65535, (256LL << 20) / ((int64_t)nx * nz))));
    const size_t per = (size_t)nx * nz;
    if (int rc = dev_grow(&e->d_slots, &e->slots_cap, (size_t)chunk)) return rc;
    if (int rc = ensure_ident(e, chunk)) return rc;
    if (int rc = dev_grow(&e->d_seeds, &e->seeds_cap, (size_t)chunk)) return rc;
    if (int rc = dev_grow(&e->d_field, &e->field_cap, (size_t)chunk * per)) return rc;
    for (int s0 = 0; s0 < ns; s0 += chunk) {
        const int n = std::min(chunk, ns - s0);
        MCB_HIPCHK(hipMemcpy(e->d_seeds, seeds + s0, n * sizeof(int64_t), hipMemcpyHostToDevice));
        MCB_HIPCHK(launch_seed_init(e->w, e->d_seeds, e->d_ident, n, e->d_slots, nullptr));
        MCB_HIPCHK(launch_end_field(e->w, e->d_slots, n, bx0, bz0, nx, nz, step, kind, e->d_field, nullptr));
        MCB_HIPCHK(hipMemcpy(out + (size_t)s0 * per, e->d_field, (size_t)n * per * sizeof(double), hipMemcpyDeviceToHost));
    }
    return MCB_OK;
}

int mcb_points_dev(McbEngine *e, const int64_t *d_seeds, int ns, const int32_t *d_pts, int np, uint8_t *d_out,
                   int32_t *d_tg, void *stream) {
    if (!e || ns < 0 || np < 0 || (ns && np && (!d_seeds || !d_pts || !d_out))) return fail(MCB_E_ARG, "bad arguments");
    if (ns == 0 || np == 0) return MCB_OK;
    std::lock_guard<std::mutex> lk(e->mu);
    return points_locked(e, d_seeds, ns, d_pts, np, d_out, d_tg, (hipStream_t)stream);
}

int mcb_points(McbEngine *e, const int64_t *seeds, int ns, const int32_t *pts, int np, uint8_t *out, int32_t *tg) {
    if (!e || ns < 0 || np < 0 || (ns && np && (!seeds || !pts || !out))) return fail(MCB_E_ARG, "bad arguments");
    if (ns == 0 || np == 0) return MCB_OK;
    const size_t n = (size_t)ns * np;
    std::lock_guard<std::mutex> lk(e->mu);
    if (int rc = dev_grow(&e->d_seeds, &e->seeds_cap, (size_t)ns)) return rc;
    if (int rc = dev_grow(&e->d_pts, &e->pts_cap, (size_t)np * 3)) return rc;
    if (int rc = dev_grow(&e->d_out, &e->out_cap, n)) return rc;
    if (tg)
        if (int rc = dev_grow(&e->d_tg, &e->tg_cap, n * 6)) return rc;
    MCB_HIPCHK(hipMemcpy(e->d_seeds, seeds, ns * sizeof(int64_t), hipMemcpyHostToDevice));
    MCB_HIPCHK(hipMemcpy(e->d_pts, pts, (size_t)np * 3 * sizeof(int32_t), hipMemcpyHostToDevice));
    if (int rc = points_locked(e, e->d_seeds, ns, e->d_pts, np, e->d_out, tg ? e->d_tg : nullptr, nullptr)) return rc;
    MCB_HIPCHK(hipMemcpy(out, e->d_out, n, hipMemcpyDeviceToHost));
    if (tg) MCB_HIPCHK(hipMemcpy(tg, e->d_tg, n * 6 * sizeof(int32_t), hipMemcpyDeviceToHost));
    return MCB_OK;
}

int mcg_host_load(McgHost *H, int version, int dim, const char *preset) {
    if (!H) return -1;
    memset(H, 0, sizeof *H);
    H->version = version;
    H->dim = dim;
    McbEngine *e = nullptr;
    const char *p = (!preset || !strcmp(preset, "normal")) ? "overworld" : preset;
    if (mcb_open(&e, version, dim, dim == MCB_OVERWORLD ? p : "overworld", nullptr, 0) != MCB_OK) return -1;
    H->impl = e;
    H->mode = e->w.host.fmt_new;
    snprintf(H->preset, sizeof H->preset, "%s", dim == MCB_OVERWORLD ? p : "-");
    return 0;
}

void mcg_host_free(McgHost *H) {
    if (!H) return;
    mcb_close((McbEngine *)H->impl);
    H->impl = nullptr;
}

// The reference's void entry points cannot fail; these can (no device, a device error).  A failure must
// not look like a result: with no handler installed it is printed and the process aborts; with one, the
// handler gets the message and the outputs hold 0xFF bytes (not a biome id) / empty masks.
static McgRefErrorFn g_ref_err_fn = nullptr;
static void *g_ref_err_ud = nullptr;

void mcg_ref_set_error_handler(McgRefErrorFn fn, void *ud) {
    g_ref_err_fn = fn;
    g_ref_err_ud = ud;
}

static void ref_failure(const char *where) {
    char msg[600];
    snprintf(msg, sizeof msg, "%s: %s", where, g_err.c_str());
    if (!g_ref_err_fn) {
        fprintf(stderr, "%s\n", msg);
        abort();
    }
    g_ref_err_fn(msg, g_ref_err_ud);
}

void mcg_ref_points(const McgHost *H, const int64_t *seeds, int ns, const int32_t *pts, int np, uint8_t *out_biome,
                    int32_t *out_tg) {
    if (!H || !H->impl) {
        fail(MCB_E_ARG, "engine not loaded");
    } else if (mcb_points((McbEngine *)H->impl, seeds, ns, pts, np, out_biome, out_tg) == MCB_OK) {
        return;
    }
    if (out_biome && ns > 0 && np > 0) memset(out_biome, 0xFF, (size_t)ns * np);
    ref_failure("mcg_ref_points");
}

void mcg_ref_tie_mask(const McgHost *H, const int32_t tg[6], uint64_t out[2]) {
    if (out) out[0] = out[1] = 0;
    if (!H || !H->impl || !tg || !out) {
        fail(MCB_E_ARG, "bad arguments");
    } else if (mcb_tie_masks((McbEngine *)H->impl, tg, 1, out) == MCB_OK) {
        return;
    }
    ref_failure("mcg_ref_tie_mask");
}

}  // extern "C"
