// mcb_internal.h - host-side internals shared by the C-ABI layer, the plans and the server.
#ifndef MCB_INTERNAL_H
#define MCB_INTERNAL_H
#include <hip/hip_runtime.h>

#include <cstdarg>

#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mcbiome.h"
#include "mcb_common.h"

namespace mcb {

int biome_count();
const char *biome_name(int id);
int biome_id(const char *name);

// Fills the world constants and the pre-order R-tree from the data directory.  0 on success.
int build_world(const std::string &root, int version, int dim, const std::string &preset, McbWorld *w,
                std::vector<McbNode> &tree, std::string &err);

// Climate cell table (mcb_cells.cpp): per cell of the breakpoint grid, the ascending pre-order
// indices of the leaves that can be the (distance, index) minimum of a target inside it.
// Cell index = (((((cT*nH + cH)*nC + cC)*nE + cE)*nW + cW)*nD + cD), n = breakpoints + 1, cell i of a
// dimension = [bp[i-1], bp[i]-1] with the outer cells bounded by +-kCellTMax (beyond every climate
// noise's bound, so only a depth or a caller's target can fall outside and walk the tree).
constexpr int kCellTMax = 1 << 17;
constexpr int kCellMaxBp = MCB_CELL_MAX_BP;
struct CellTable {
    int nb[6] = {0, 0, 0, 0, 0, 0};    // breakpoints per tree dimension (0 T, 1 H, 2 C, 3 E, 4 depth, 5 W)
    int off[6] = {0, 0, 0, 0, 0, 0};   // first breakpoint of each dimension in bp
    std::vector<int32_t> bp;
    std::vector<uint32_t> entry;       // (first << 8) | count; count 0 = walk the tree
    std::vector<uint16_t> list;
    double build_ms = 0.0;
};
// strict: a leaf is left out of a cell only when another leaf is strictly nearer everywhere in it, so the
// list also holds every leaf that can tie at the least distance (the plugin's tie test needs them)
int build_cells(const std::vector<McbNode> &tree, CellTable &ct, std::string &err, bool strict = false);
std::shared_ptr<const CellTable> cells_for_tree(const std::vector<McbNode> &tree, std::string &err, bool strict = false,
                                               double max_cells = 0.0);

// One entry of a biome parameter list (params/<V>/<dim>.tsv): quantized intervals of temperature,
// humidity, continentalness, erosion, depth, weirdness and the offset (lo = hi), and the biome id.
struct ParamEntry {
    int64_t lo[7], hi[7];
    int biome;
};
// Climate.RTree of a parameter list in device pre-order (mcb_rtree.cpp); exclusive_hi: 26.4-snapshot-2's
// upper bounds, lowered by one after the build
int build_tree(const std::vector<ParamEntry> &list, int per_node, bool exclusive_hi, std::vector<McbNode> &out,
               std::string &err);

// BiomeManager.obfuscateSeed (SHA-256 of the seed's little-endian bytes, first 8 digest bytes LE)
int64_t hashed_seed(int64_t seed);

// Device copy of one world.
struct DevWorld {
    McbWorld host;
    McbWorld *d_world = nullptr;
    McbNode *d_tree = nullptr;
    uint32_t *d_cells = nullptr;      // Overworld cell table (null elsewhere): entries, see cell_best
    McbNode *d_cell_recs = nullptr;   // the listed leaves, copied in list order
    // per cell of the five dimensions other than depth: the union of the lists of its depth cells,
    // (first << 8) | count into d_cell_urecs (count 0: no union, each layer scans its own list)
    uint32_t *d_cell_u = nullptr;
    McbNode *d_cell_urecs = nullptr;

    double cells_build_ms = 0.0;
    int version = 0, dim = 0;
    std::string preset;
};

// error reporting: sets the thread-local message returned by mcb_last_error, returns `code`
int fail(int code, const char *fmt, ...);
int failv(int code, const char *fmt, va_list ap);
std::string default_data_dir();

#define MCB_HIPCHK(x)                                                                                    \
    do {                                                                                                 \
        hipError_t e_ = (x);                                                                             \
        if (e_ != hipSuccess) return ::mcb::fail(MCB_E_DEVICE, "%s: %s", #x, hipGetErrorString(e_));     \
    } while (0)

template <class T>
int dev_grow(T **p, size_t *cap, size_t n) {
    if (n <= *cap) return 0;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    if (hipMalloc((void **)p, n * sizeof(T)) != hipSuccess) return fail(MCB_E_NOMEM, "hipMalloc %zu bytes", n * sizeof(T));
    *cap = n;
    return 0;
}

// ---- kernel launchers (mcb_kernels.hip); all asynchronous on `stream` ----
hipError_t launch_seed_init(const DevWorld &w, const int64_t *d_seeds, const int32_t *d_slot_of, int n,
                            McbSlot *d_slots, hipStream_t stream);
// tile = tile_w x (256/tile_w)*cpl columns, cpl (1/2/4) columns per lane; max_ny = most layers of any request
// nowrap: every octave input of these requests stays inside +-2^24 (wrap_free), so the 26.3 kernel may
// leave out PerlinNoise.wrap's range test
hipError_t launch_grid(const DevWorld &w, const McbSlot *d_slots, const McbRequest *d_reqs, const McbTile *d_tiles,
                       int ntiles, uint8_t *d_out, int tile_w, int cpl, int max_ny, int grid_blocks, bool nowrap,
                       hipStream_t stream);
hipError_t launch_points(const DevWorld &w, const McbSlot *d_slots, const int32_t *d_slot_of, int ns,
                         const int32_t *d_pts, int np, uint8_t *d_out, int32_t *d_tg, hipStream_t stream);
// biomes of n quantized climate targets (6 int32 each: t,h,c,e,d,w); Overworld and Nether only
hipError_t launch_targets(const DevWorld &w, const int32_t *d_tg, int n, uint8_t *d_out, hipStream_t stream);
hipError_t launch_tie_masks(const DevWorld &w, const int32_t *d_tg, int n, uint64_t *d_out, hipStream_t stream);
// raw climate floats (6 per sample) at block positions, slots[s] = seed s (mcb_climate.hip)
hipError_t launch_climate_raw(const DevWorld &w, const McbSlot *d_slots, int ns, const int32_t *d_xyz, int np, float *d_out,
                              hipStream_t stream);
// one NormalNoise of the world (index into McbWorld.noise) at double positions, slots[s] = seed s
hipError_t launch_noise_values(const DevWorld &w, const McbSlot *d_slots, int ns, int noise, const double *d_xyz, int np,
                               double *d_out, hipStream_t stream);
// End island field of ns seeds (slots 0..ns-1) over an nx x nz block grid; kind 0 height, 1 erosion
hipError_t launch_end_field(const DevWorld &w, const McbSlot *d_slots, int ns, int bx0, int bz0, int nx, int nz, int step,
                            int kind, double *d_out, hipStream_t stream);
hipError_t launch_zoom(const McbRequest *d_zreqs, const McbRequest *d_qreqs, const McbTile *d_tiles, int ntiles,
                       const uint8_t *d_quart, uint8_t *d_out, hipStream_t stream);

// ---- request lowering shared by plans and the server ----
// Appends the kernel descriptors of one region: a direct quart grid (g) or, for zoom, the quart layers
// (q) feeding a zoom request (z).  quart_off accumulates the zoom staging bytes.
struct Lowered {
    std::vector<McbRequest> g, q, z;
    std::vector<McbTile> gt, qt, zt;
    int64_t quart_bytes = 0, columns = 0, samples = 0;
    int max_ny_g = 1, max_ny_q = 1;
};
int check_region(const McbRegion &r);
// True when no octave input of the quart-grid requests' climate columns can reach +-2^24 (so PerlinNoise.wrap,
// engine/mc_noise.h:86-89, is the identity on every one): the largest quart coordinate, widened by the most
// the offset noise can shift it (4 x twice the summed amplitudes of both stacks, 26.1/26.2's amplitudes
// times their level and NormalNoise value factors; every octave value is below 2 in magnitude), times the
// largest frequency of any octave (and both stacks' input factor), stays below 2^23.
bool wrap_free(const McbWorld &w, const std::vector<McbRequest> &reqs);
void lower_region(const McbRegion &r, int slot, int tile_w, int cpl, Lowered &L);

// dst[0..bytes) (host memory, pageable or not) = src[0..bytes) (device memory), after the work queued on st:
// blocks staged through pinned buffers on st while host threads copy the previous block out; returns when
// the bytes are in dst.  Copies below 1 MiB, and copies into pinned host or device memory, are one
// hipMemcpyAsync + stream synchronisation.
hipError_t copy_to_host(void *dst, const void *src, size_t bytes, hipStream_t st);

}  // namespace mcb

// Engine object behind the C ABI's opaque McbEngine.
struct McbEngine {
    mcb::DevWorld w;
    int device = 0;
    std::mutex mu;
    // one-off regions (mcb_region): a plan group rebuilt per call in grow-only buffers (mcb_plan.cpp),
    // guarded by region_mu; freed by region_free at close
    std::mutex region_mu;
    void *region_cache = nullptr;
    void (*region_free)(void *) = nullptr;
    // per-call scratch (grow-only), guarded by mu
    McbSlot *d_slots = nullptr;
    size_t slots_cap = 0;
    int32_t *d_ident = nullptr;
    size_t ident_cap = 0;
    int64_t *d_seeds = nullptr;
    size_t seeds_cap = 0;
    int32_t *d_pts = nullptr;
    size_t pts_cap = 0;
    uint8_t *d_out = nullptr;
    size_t out_cap = 0;
    int32_t *d_tg = nullptr;
    size_t tg_cap = 0;
    size_t cell_entries = 0, cell_list_len = 0;
    double *d_field = nullptr;
    size_t field_cap = 0;
    uint64_t *d_masks = nullptr;
    size_t masks_cap = 0;
    float *d_raw = nullptr;                // raw climate floats
    size_t raw_cap = 0;
    double *d_xyzd = nullptr;              // double positions of noise queries
    size_t xyzd_cap = 0;
    // host copies kept for the seed search's prefix tables
    std::shared_ptr<const mcb::CellTable> cells;
    std::vector<McbNode> tree_host;
    // seed search (mcb_search.cpp), grow-only
    uint4 *d_feas = nullptr;               // prefix biome sets
    uint32_t *d_coord_lut = nullptr;
    int feas_off[4] = {0, 0, 0, 0};
    bool feas_ok = false;                  // pruning allowed (noise bounds inside the table)
    McbObs *d_obs = nullptr;
    unsigned long long *d_hist = nullptr;
    uint64_t *d_next = nullptr;
    uint32_t *d_nfound = nullptr;
    int64_t *d_found = nullptr;
    size_t found_cap = 0;
    int64_t *d_list = nullptr;
    size_t list_cap = 0;
};

struct McgBiomeWorld;
namespace mcb {
// identity slot map (slot i = seed i) of at least n entries in e->d_ident
int ensure_ident(McbEngine *e, int n);
// renumber the biome ids of a world made by mcb_plugin_world (leaves, End ids, fixed biome): id -> map[id]
void plugin_world_renumber(McgBiomeWorld *W, const uint8_t map[256]);
}  // namespace mcb

#endif
