// mcgen_api.cpp - the biome part of libmcgen's public C ABI (include/mcgen.h; reference:
// libmcgen/include/mcgen.h:1-172, libmcgen/src/gpu_bridge.c:163-240, libmcgen/src/biome.c).
//
// A McWorld is one vanilla world (version, dimension, preset, climate seed) compiled once into the GPU
// plugin's device form (mcb_plugin_world -> the plugin compiler) with tie chains uncapped, as libmcgen's
// CPU path has no cap; mcgen_biome_at / mcgen_biome_grid / mcgen_gpu_biome_points run the plugin's grid
// and points kernels on it and mcgen_generate_region(MC_STAGE_BIOMES) its chunk kernel.  Everything else
// libmcgen does (terrain, surface, carvers, features, structures, block states, tweaks, schedules) is
// outside this library: those entry points report unsupported or empty.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "../../include/mcgen.h"
#include "../../include/mcgpu.h"
#include "mcb_internal.h"
#include "mcg_plugin.h"

namespace {
// A version's biome registry as libmcgen numbers it: the biome files of the version's pack, sorted by
// name (libmcgen/src/pack.c:134-146 with list_dir_recursive's qsort, libmcgen/src/util.c:169-173).  The
// vanilla files are the biomes the version's Overworld and Nether parameter lists place, the five End
// biomes and the_void (26.1: 65, 26.2: 66, 26.3: 67).  from_engine maps the engine's ids to these.
struct Registry {
    std::vector<std::string> names;       // "minecraft:<name>", id order
    uint8_t from_engine[256];
};
const char *const kUnplaced[6] = {"the_end", "end_highlands", "end_midlands", "small_end_islands", "end_barrens", "the_void"};

int load_registry(int version, Registry &R, std::string &err) {
    std::set<std::string> names(kUnplaced, kUnplaced + 6);
    for (const char *dim : {"overworld", "nether"}) {
        const std::string path = mcb::default_data_dir() + "/params/" + mcg_version_name(version) + "/" + dim + ".tsv";
        std::ifstream f(path);
        if (!f) { err = "cannot read " + path; return MCGEN_E_IO; }
        std::string line;
        while (std::getline(f, line)) {
            if (line.empty() || line[0] == '#') continue;
            std::string b = line.substr(line.rfind('\t') + 1);
            while (!b.empty() && (b.back() == '\r' || b.back() == ' ')) b.pop_back();
            if (b.compare(0, 10, "minecraft:") == 0) b = b.substr(10);
            names.insert(b);
        }
    }
    for (auto &n : names) R.names.push_back("minecraft:" + n);          // std::set orders as strcmp
    memset(R.from_engine, 255, sizeof R.from_engine);
    for (int i = 0; i < mcg_biome_count() && i < 256; i++) {
        const std::string full = std::string("minecraft:") + mcg_biome_name(i);
        for (size_t j = 0; j < R.names.size(); j++)
            if (R.names[j] == full) R.from_engine[i] = (uint8_t)j;
    }
    return MCGEN_OK;
}

const Registry *registry(int version, std::string &err) {
    static std::mutex mu;
    static std::map<int, Registry> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(version);
    if (it != cache.end()) return &it->second;
    Registry R;
    if (load_registry(version, R, err)) return nullptr;
    return &(cache[version] = std::move(R));
}
}  // namespace

struct McGen {
    int version = 0;
    std::string vname;
    const Registry *reg = nullptr;
};

struct McWorld {
    McGen *g = nullptr;
    int dim = 0;
    std::string preset;
    int64_t climate_seed = 0;
    int min_y = 0, height = 0, sea_level = 0;
    void *h = nullptr;                       // compiled plugin world (uncapped tie chains)
};

struct McRegion {
    McRegionInfo info;
    int qh = 0;
    std::vector<uint8_t> biomes;             // chunk (cz - cz0) * nx + (cx - cx0), qh*16 bytes each
};

namespace {

const char *const kDims[3] = {"minecraft:overworld", "minecraft:the_nether", "minecraft:the_end"};
const char *const kOwPresets[3] = {"normal", "large_biomes", "amplified"};
int g_compute = MCGEN_COMPUTE_AUTO;

void put_err(char *err, size_t errlen, const char *fmt, const char *a = "") {
    if (err && errlen) snprintf(err, errlen, fmt, a);
}
int dim_of(const char *d) {
    if (!d) return -1;
    const char *s = strncmp(d, "minecraft:", 10) == 0 ? d + 10 : d;
    if (!strcmp(s, "overworld")) return MCB_OVERWORLD;
    if (!strcmp(s, "the_nether") || !strcmp(s, "nether")) return MCB_NETHER;
    if (!strcmp(s, "the_end") || !strcmp(s, "end")) return MCB_END;
    return -1;
}
int grid_code(const McWorld *w, int x0, int z0, int nx, int nz, int step, int y, uint8_t *out) {
    if (!w || !w->h || nx <= 0 || nz <= 0 || step <= 0 || !out) return MCGEN_E_ARG;
    int bad = 0;
    if (mcgpu_biome_grid(w->h, x0, z0, nx, nz, step, y, out, &bad)) return MCGEN_E_INTERNAL;
    return bad ? MCGEN_E_INTERNAL : MCGEN_OK;        // a non-finite climate: not produced by vanilla worlds
}

}  // namespace

extern "C" {

const char *mcgen_version(void) { return "libmcgen-compatible biome stage (libmcbiome_hip) abi 1"; }
int mcgen_abi_version(void) { return MCGEN_ABI_VERSION; }

int mcgen_open(const char *pack_dir, const char *version, McGen **out, char *err, size_t errlen) {
    (void)pack_dir;                                  // the built-in vanilla tables are used
    if (!out || !version) { put_err(err, errlen, "bad arguments"); return MCGEN_E_ARG; }
    *out = nullptr;
    const int v = mcg_version_from_name(version);
    if (v < 0 || v > MCB_V26_3) { put_err(err, errlen, "unsupported version %s (26.1, 26.2, 26.3)", version); return MCGEN_E_VERSION; }
    std::string rerr;
    const Registry *reg = registry(v, rerr);
    if (!reg) { put_err(err, errlen, "%s", rerr.c_str()); return MCGEN_E_IO; }
    McGen *g = new McGen();
    g->version = v;
    g->vname = version;
    g->reg = reg;
    *out = g;
    return MCGEN_OK;
}
void mcgen_close(McGen *g) { delete g; }

int mcgen_dimension_count(const McGen *g) { return g ? 3 : 0; }
const char *mcgen_dimension_name(const McGen *g, int i) { return g && i >= 0 && i < 3 ? kDims[i] : nullptr; }
int mcgen_preset_count(const McGen *g, const char *dimension) {
    const int d = dim_of(dimension);
    return !g || d < 0 ? 0 : (d == MCB_OVERWORLD ? 3 : 1);
}
const char *mcgen_preset_name(const McGen *g, const char *dimension, int i) {
    const int n = mcgen_preset_count(g, dimension);
    if (i < 0 || i >= n) return nullptr;
    return dim_of(dimension) == MCB_OVERWORLD ? kOwPresets[i] : "normal";
}
int mcgen_block_state_count(const McGen *g) { (void)g; return 0; }
const char *mcgen_block_state_name(const McGen *g, int id) { (void)g; (void)id; return nullptr; }
int mcgen_block_state_from_name(const McGen *g, const char *name) { (void)g; (void)name; return -1; }
int mcgen_biome_count(const McGen *g) { return g ? (int)g->reg->names.size() : 0; }
const char *mcgen_biome_name(const McGen *g, int id) {
    return g && id >= 0 && id < (int)g->reg->names.size() ? g->reg->names[id].c_str() : nullptr;
}
int mcgen_tweak_count(const McGen *g) { (void)g; return 0; }
const McTweakInfo *mcgen_tweak_info(const McGen *g, int i) { (void)g; (void)i; return nullptr; }

int mcgen_world_new(McGen *g, const char *dimension, const char *preset, const McSeeds *seeds,
                    const McTweakValue *tweaks, int ntweaks, McWorld **out, char *err, size_t errlen) {
    (void)tweaks;
    if (!out) { put_err(err, errlen, "bad arguments"); return MCGEN_E_ARG; }
    *out = nullptr;
    const int d = dim_of(dimension);
    if (!g || !seeds || d < 0) { put_err(err, errlen, "bad arguments (dimension %s)", dimension ? dimension : "null"); return MCGEN_E_ARG; }
    if (ntweaks > 0) { put_err(err, errlen, "world tweaks are not supported"); return MCGEN_E_UNSUPPORTED; }
    const char *p = (!preset || !*preset || !strcmp(preset, "normal")) ? "overworld" : preset;
    if (d == MCB_OVERWORLD && strcmp(p, "overworld") && strcmp(p, "large_biomes") && strcmp(p, "amplified")) {
        put_err(err, errlen, "unknown preset %s", preset);
        return MCGEN_E_ARG;
    }
    McbEngine *e = nullptr;
    if (mcb_open(&e, g->version, d, d == MCB_OVERWORLD ? p : "overworld", nullptr, 0) != MCB_OK) {
        put_err(err, errlen, "%s", mcb_last_error());
        return MCGEN_E_DATA;
    }
    McgBiomeWorld *pw = nullptr;
    if (mcb_plugin_world(e, seeds->climate, &pw) != MCB_OK) {
        put_err(err, errlen, "%s", mcb_last_error());
        mcb_close(e);
        return MCGEN_E_INTERNAL;
    }
    mcb::plugin_world_renumber(pw, g->reg->from_engine);     // libmcgen's ids of this version
    char perr[512] = {0};
    void *h = mcb::plug::world_new(pw, 1 << 30, perr, sizeof perr);
    mcb_plugin_world_free(pw);
    mcb_close(e);
    if (!h) { put_err(err, errlen, "%s", perr); return MCGEN_E_INTERNAL; }
    McWorld *w = new McWorld();
    w->g = g;
    w->dim = d;
    w->preset = p;
    w->climate_seed = seeds->climate;
    w->h = h;
    w->min_y = d == MCB_OVERWORLD ? -64 : 0;
    w->height = d == MCB_OVERWORLD ? 384 : 256;
    w->sea_level = d == MCB_OVERWORLD ? 63 : (d == MCB_NETHER ? 32 : 0);
    *out = w;
    return MCGEN_OK;
}
void mcgen_world_free(McWorld *w) {
    if (!w) return;
    mcgpu_biome_free(w->h);
    delete w;
}
int mcgen_world_set_schedule(McWorld *w, const char *path, char *err, size_t errlen) {
    if (!w) return MCGEN_E_ARG;
    if (!path || !*path) return MCGEN_OK;            // no schedule: the only state here
    put_err(err, errlen, "schedules order decoration steps, which this library does not run");
    return MCGEN_E_UNSUPPORTED;
}
int mcgen_world_min_y(const McWorld *w) { return w ? w->min_y : 0; }
int mcgen_world_height(const McWorld *w) { return w ? w->height : 0; }
int mcgen_world_sea_level(const McWorld *w) { return w ? w->sea_level : 0; }

int mcgen_biome_at(const McWorld *w, int x, int y, int z) {
    if (!w || !w->h) return -1;
    const int xyz[3] = {x, y, z};
    uint8_t b = 0;
    int bad = 0;
    if (mcgpu_biome_points(w->h, 1, xyz, &b, &bad) || bad) return -1;
    return b;
}
int mcgen_biome_grid(const McWorld *w, int x0, int z0, int nx, int nz, int step, int y, uint8_t *out) {
    return grid_code(w, x0, z0, nx, nz, step, y, out);
}

int mcgen_generate_region(McWorld *w, int cx0, int cz0, int nx, int nz, uint32_t stages, int threads, McProgressFn cb,
                          void *ud, McRegion **out, char *err, size_t errlen) {
    (void)threads;
    if (!out) return MCGEN_E_ARG;
    *out = nullptr;
    // at most 65,536 chunks per call (100 MB of Overworld biomes on the host)
    if (!w || !w->h || nx <= 0 || nz <= 0 || (int64_t)nx * nz > (1 << 16)) { put_err(err, errlen, "bad arguments"); return MCGEN_E_ARG; }
    if (stages & ~(uint32_t)MC_STAGE_BIOMES) {
        put_err(err, errlen, "only MC_STAGE_BIOMES is computed by this library");
        return MCGEN_E_UNSUPPORTED;
    }
    std::unique_ptr<McRegion> r(new McRegion());
    r->info = {cx0, cz0, nx, nz, w->min_y, w->height};
    r->qh = w->height / 4;
    const int n = nx * nz;
    std::vector<int32_t> cxz((size_t)n * 2);
    for (int iz = 0; iz < nz; iz++)
        for (int ix = 0; ix < nx; ix++) {
            cxz[2 * ((size_t)iz * nx + ix)] = cx0 + ix;
            cxz[2 * ((size_t)iz * nx + ix) + 1] = cz0 + iz;
        }
    if (cb && cb(ud, 0.0, "biomes")) return MCGEN_E_CANCEL;
    r->biomes.resize((size_t)n * r->qh * 16);
    int bad = 0;
    if (mcb::plug::world_chunks(w->h, n, cxz.data(), r->biomes.data(), &bad)) {
        put_err(err, errlen, "%s", mcgpu_last_error());
        return MCGEN_E_INTERNAL;
    }
    if (bad) { put_err(err, errlen, "non-finite climate"); return MCGEN_E_INTERNAL; }
    if (cb && cb(ud, 1.0, "biomes")) return MCGEN_E_CANCEL;
    *out = r.release();
    return MCGEN_OK;
}
void mcgen_region_free(McRegion *r) { delete r; }
void mcgen_region_info(const McRegion *r, McRegionInfo *info) {
    if (r && info) *info = r->info;
}
uint16_t *mcgen_region_blocks(McRegion *r, int cx, int cz) { (void)r; (void)cx; (void)cz; return nullptr; }
uint8_t *mcgen_region_biomes(McRegion *r, int cx, int cz) {
    if (!r) return nullptr;
    const int ix = cx - r->info.cx0, iz = cz - r->info.cz0;
    if (ix < 0 || iz < 0 || ix >= r->info.nx || iz >= r->info.nz) return nullptr;
    return r->biomes.data() + ((size_t)iz * r->info.nx + ix) * r->qh * 16;
}
int16_t *mcgen_region_heightmap(McRegion *r, int cx, int cz, int kind) {
    (void)r; (void)cx; (void)cz; (void)kind;
    return nullptr;
}

int mcgen_gpu_set_features(int mask) { (void)mask; return MCGEN_OK; }
int mcgen_gpu_get_features(void) { return MCGEN_GPU_BIOMES; }
int mcgen_gpu_set_library_path(const char *path) { (void)path; return MCGEN_OK; }
int mcgen_gpu_device_count(void) { return mcgpu_device_count(); }
int mcgen_gpu_device_info(int index, char *name, size_t namelen, int *cc_major, int *cc_minor, size_t *mem_mb) {
    return mcgpu_device_info(index, name, namelen, cc_major, cc_minor, mem_mb) ? MCGEN_E_ARG : MCGEN_OK;
}
int mcgen_gpu_set_compute(int mode, int device) {
    if (mode < MCGEN_COMPUTE_CPU || mode > MCGEN_COMPUTE_AUTO) return MCGEN_E_ARG;
    if (mode == MCGEN_COMPUTE_CPU) return MCGEN_E_UNSUPPORTED;        // no CPU path in this library
    if (device >= 0 && mcgpu_select(device)) return MCGEN_E_UNSUPPORTED;
    g_compute = mode;
    return MCGEN_OK;
}
int mcgen_gpu_get_compute(void) { return g_compute; }
int mcgen_gpu_status(char *buf, size_t buflen) {
    char name[128] = "-";
    int ma = 0, mi = 0;
    size_t mem = 0;
    const int n = mcgpu_device_count();
    if (n > 0) mcgpu_device_info(mcgpu_selected() >= 0 ? mcgpu_selected() : 0, name, sizeof name, &ma, &mi, &mem);
    if (buf && buflen)
        snprintf(buf, buflen, "mode: %s\nstate: %s\nfeatures: biomes\ndevice: %s (gfx%d%d, %zu MB)\nlibrary: libmcbiome_hip.so",
                 g_compute == MCGEN_COMPUTE_GPU ? "GPU" : "Auto", n > 0 ? "GPU ready" : "no HIP device", name, ma, mi, mem);
    return n > 0 ? 1 : 0;
}
// The library has no CPU path to compare with; the self-test checks the two device paths against each
// other instead: a block grid through the cell lattice and the same blocks as independent points.
int mcgen_gpu_selftest(McGen *g, int npoints, char *report, size_t replen) {
    if (!g) return MCGEN_E_ARG;
    if (npoints <= 0) npoints = 256;
    int side = 1;
    while (side * side < npoints) side++;
    std::string rep;
    int mism = 0;
    for (int d = 0; d < 3; d++) {
        const McSeeds s = mcgen_seeds_unified(0x5EED + d);
        McWorld *w = nullptr;
        char err[256] = {0};
        if (mcgen_world_new(g, kDims[d], "normal", &s, nullptr, 0, &w, err, sizeof err) != MCGEN_OK) {
            put_err(report, replen, "world: %s", err);
            return MCGEN_E_INTERNAL;
        }
        std::vector<uint8_t> grid((size_t)side * side), pts((size_t)side * side);
        std::vector<int32_t> xyz((size_t)side * side * 3);
        const int x0 = -37 * side, z0 = 53 * side, y = w->min_y + 70;
        for (int i = 0; i < side * side; i++) {
            xyz[3 * i] = x0 + (i % side);
            xyz[3 * i + 1] = y;
            xyz[3 * i + 2] = z0 + i / side;
        }
        int bad = 0;
        const bool ok = grid_code(w, x0, z0, side, side, 1, y, grid.data()) == MCGEN_OK &&
                        mcgpu_biome_points(w->h, side * side, xyz.data(), pts.data(), &bad) == 0 && !bad;
        int m = 0;
        for (int i = 0; ok && i < side * side; i++) m += grid[i] != pts[i];
        mism += ok ? m : side * side;
        rep += std::string(kDims[d]) + (ok ? ": " + std::to_string(m) + " mismatches\n" : ": device error\n");
        mcgen_world_free(w);
    }
    put_err(report, replen, "%s", rep.c_str());
    return mism ? MCGEN_E_INTERNAL : MCGEN_OK;
}
int mcgen_gpu_biome_grid(const McWorld *w, int x0, int z0, int nx, int nz, int step, int y, uint8_t *out) {
    return grid_code(w, x0, z0, nx, nz, step, y, out);
}
int mcgen_biome_grid_cpu(const McWorld *w, int x0, int z0, int nx, int nz, int step, int y, uint8_t *out) {
    (void)w; (void)x0; (void)z0; (void)nx; (void)nz; (void)step; (void)y; (void)out;
    return MCGEN_E_UNSUPPORTED;                      // no CPU path in this library
}
int mcgen_gpu_biome_points(const McWorld *w, int n, const int *xyz, uint8_t *out, int *n_cpu) {
    if (!w || !w->h || n < 0 || (n && (!xyz || !out))) return MCGEN_E_ARG;
    if (n_cpu) *n_cpu = 0;
    int bad = 0;
    if (mcgpu_biome_points(w->h, n, xyz, out, &bad)) return MCGEN_E_INTERNAL;
    return bad ? MCGEN_E_INTERNAL : MCGEN_OK;
}

}  // extern "C"
