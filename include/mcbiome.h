/* mcbiome.h - C ABI of the MI355X (gfx950) biome engine: libmcbiome_hip.so
 *
 * Plain C types only (pointers, sizes, 64-bit seeds); no HIP or torch types.  Every entry point
 * returns 0 on success and a nonzero MCB_E_* code on failure, with the message available from
 * mcb_last_error() (thread-local).  The library never falls back to a CPU path: without a usable
 * HIP device every compute entry point fails with MCB_E_DEVICE.
 *
 * Biome ids are u8 indices into the 67-name alphabetical table (mcg_biome_name), the ids of the
 * reference engine (engine/gen/mc_biome_names.h:4-5).  Versions: 0 = 26.1, 1 = 26.2, 2 = 26.3,
 * 3 = 26.4-snapshot-2 (engine/mc_common.h:35-38).
 * Dimensions: 0 = overworld, 1 = nether, 2 = end.  Search semantics are the reference CPU engine's:
 * each sample's R-tree search starts from no candidate (engine/mc_biomes.h:58-62).
 *
 * Data: climate-<V>.txt, params/<V>/{overworld,nether}.tsv and rtree_<V>_{overworld,the_nether}.json
 * are read from `data_dir`, or when that is NULL from $MCB_DATA, else $MCGEN_ROOT/data (the
 * reference engine's convention, engine/mc_data.h:19), else the data/ folder beside the library.
 */
#ifndef MCBIOME_H
#define MCBIOME_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum { MCB_OK = 0, MCB_E_ARG = 1, MCB_E_IO = 2, MCB_E_DATA = 3, MCB_E_DEVICE = 4, MCB_E_NOMEM = 5, MCB_E_LIMIT = 6 };

/* ================================================================================================
 * Drop-in for the reference CPU engine API, engine/cuda/mcg_host.h:17-45.  Same names, argument
 * meaning, output layout and return codes; the work runs on the HIP device.  McgHost keeps the
 * reference's public fields (version, dim, mode, preset); the engine state behind it is opaque.
 * ================================================================================================ */
typedef struct McgHost {
    int version;        /* 0 = 26.1, 1 = 26.2, 2 = 26.3 */
    int dim;            /* 0 overworld, 1 nether, 2 end */
    int mode;           /* 0 = double noise (26.1/26.2), 1 = float noise (26.3) */
    char preset[16];    /* "overworld" | "amplified" | "large_biomes" | "-" */
    void *impl;         /* McbEngine* */
} McgHost;

/* replaces mcg_host_load (engine/cuda/mcg_host.h:34 / mcg_host.c:8-27): 0 = ok, -1 = failure */
int mcg_host_load(McgHost *H, int version, int dim, const char *preset);
/* replaces mcg_host_free (engine/cuda/mcg_host.h:35) */
void mcg_host_free(McgHost *H);
/* replace mcg_version_from_name / mcg_version_name / mcg_biome_count / mcg_biome_name / mcg_biome_id
 * (engine/cuda/mcg_host.h:37-42) */
int mcg_version_from_name(const char *s);
const char *mcg_version_name(int v);
int mcg_biome_count(void);
const char *mcg_biome_name(int id);
int mcg_biome_id(const char *name);
/* replaces mcg_ref_points (engine/cuda/mcg_host.h:47, mcg_host.c:36-83): host buffers, synchronous.
 * pts = np (qx,qy,qz) quart triples shared by all seeds; out_biome[s*np+p];
 * out_tg optional: 6 i32 per sample (t,h,c,e,d,w; Nether t,h,0,0,0,0; End zeros). */
void mcg_ref_points(const McgHost *H, const int64_t *seeds, int ns, const int32_t *pts, int np, uint8_t *out_biome,
                    int32_t *out_tg);
/* replaces mcg_ref_tie_mask (engine/cuda/mcg_host.h:51, mcg_host.c:85-96): out = mask of the biome ids of
 * every leaf at the least distance from the quantized target tg = {t,h,c,e,d,w} (bit b of out[b >> 6]),
 * by brute force over all leaves with the reference's 64-bit integer distances.  Overworld and Nether. */
void mcg_ref_tie_mask(const McgHost *H, const int32_t tg[6], uint64_t out[2]);
/* The reference's mcg_ref_points / mcg_ref_tie_mask return void and never fail; these run on the device and
 * can (no device, a device error).  By default a failure prints its message and calls abort(), so that it
 * can never pass for a result.  A caller that wants control installs a handler: it receives the message,
 * the call returns, and its outputs hold 0xFF bytes (not a biome id) / zero masks.  No counterpart in the
 * reference.  fn = NULL restores the default. */
typedef void (*McgRefErrorFn)(const char *message, void *user);
void mcg_ref_set_error_handler(McgRefErrorFn fn, void *user);

/* ================================================================================================
 * Native engine API (batched, device buffers).  Counterparts: gb_open / gb_points_gpu of
 * crack/src/gpu_biomes/gpu_biomes.h:22,30, and the per-call mcgpu_biome_grid of
 * libmcgen/gpu/mcgpu_biome.cu:203-249 that the batched plans replace.
 * ================================================================================================ */
typedef struct McbEngine McbEngine;
typedef struct McbPlan McbPlan;

const char *mcb_last_error(void);
int mcb_device_count(void);

/* One (version, dimension, preset) world on the current HIP device.  max_slots = seed-state slots
 * kept resident on the device (15,472 bytes each; 0 = default 8192). */
int mcb_open(McbEngine **out, int version, int dim, const char *preset, const char *data_dir, int max_slots);
/* The biome R-tree of a version's Overworld or Nether as the engine searches it, built on the host from
 * the parameter list (Climate.RTree.create, engine/mc_biomes.h:64-212 as loaded by engine/mc_data.h:153-183)
 * in DFS pre-order, 16 int32 per node: lo/hi of temperature, humidity, continentalness, erosion, depth,
 * weirdness (upper bounds lowered by one for 26.4-snapshot-2), the squared distance of the offset interval
 * to 0, the index just past the node's subtree, the leaf's biome id (-1 for a subtree), 0.  Returns the
 * node count (out may be NULL to ask for it), or a negative MCB_E_* code.  Needs no device. */
int mcb_tree(int version, int dim, const char *data_dir, int32_t *out, int cap);
void mcb_close(McbEngine *e);
int mcb_engine_info(const McbEngine *e, int *version, int *dim, int *mode, int *tree_nodes);
/* Overworld climate cell table that replaces the per-sample R-tree search of engine/mc_biomes.h:31-62
 * (same (distance, pre-order index) result): cells, total list length and host build time (built once
 * per tree per process).  Zero cells for the Nether and the End, which keep the tree walk. */
int mcb_engine_cells(const McbEngine *e, int64_t *cells, int64_t *list_len, double *build_ms);

/* Biomes of n quantized climate targets {t,h,c,e,d,w} (6 int32 each, the out_tg layout of
 * mcg_ref_points), host buffers, synchronous: the search step alone (engine/mc_biomes.h:31-62; the
 * quantized counterpart of libmcgen's world_biome_from_climate, libmcgen/src/mcgen_internal.h:166).
 * Overworld and Nether; |target| <= MCB_TARGET_MAX (distances stay exact in double). */
#define MCB_TARGET_MAX (1 << 24)
int mcb_targets(McbEngine *e, const int32_t *tg6, int n, uint8_t *out_biome);
/* Tie masks of n quantized targets (mcg_ref_tie_mask batched): out_masks[2i], out_masks[2i+1] = biome-id
 * mask of every least-distance leaf of target i.  Host buffers, synchronous; any int32 targets. */
int mcb_tie_masks(McbEngine *e, const int32_t *tg6, int n, uint64_t *out_masks);

/* Raw climate (engine-header entry mc_climate_overworld_raw, engine/mc_climate.h:113-164), batched: for each
 * seed s and block position p = (x, y, z) in xyz, out6[(s*np + p)*6 .. +6) = the sampler's floats
 * {temperature, humidity, continentalness, erosion, depth, weirdness} (26.1/26.2: the double values cast
 * to float, as the reference returns them).  Nether: the two legacy NormalNoises at (x*0.25, y*0.0, z*0.25)
 * as floats (engine/cuda/mcg_host.c:63-74) in slots 0, 1 and zeros in 2..5.  Host buffers, synchronous. */
int mcb_climate_raw(McbEngine *e, const int64_t *seeds, int ns, const int32_t *xyz, int np, float *out6);
/* NormalNoise values (engine-header entries normal_get_d / normal_get_f, engine/mc_noise.h:286-324) of one of
 * the world's climate noises, seeded as the world seeds it, at np double positions (x, y, z) for each seed:
 * out[s*np + p] (26.3 float values widened to double).  noise: Overworld 0 offset, 1 temperature,
 * 2 vegetation, 3 continentalness, 4 erosion, 5 ridge (Xoroshiro, engine/mc_climate.h:97-106); Nether 0
 * temperature, 1 vegetation (legacy Random(seed), Random(seed + 1)).  Host buffers, synchronous. */
int mcb_noise_values(McbEngine *e, int noise, const int64_t *seeds, int ns, const double *xyz, int np, double *out);

/* End island field (engine/mc_end.h:66-106) over an nx x nz block grid of every seed, blocks
 * (bx0 + ix*step, bz0 + iz*step): kind 0 = getHeightValue at section (bx/8, bz/8) (C division),
 * kind 1 = the End's erosion value at the block (what TheEndBiomeSource reads; the terrain's End
 * island density term).  out_host[s*nx*nz + iz*nx + ix] as double (26.3 values are floats).  End
 * engines only; host buffers, synchronous. */
int mcb_end_field(McbEngine *e, const int64_t *seeds, int ns, int bx0, int bz0, int nx, int nz, int step, int kind,
                  double *out_host);

/* ================================================================================================
 * Seed search (crack/src/gpu_biomes/gpu_biomes.h:55-70, gb_search): which candidate seeds show every
 * observed biome.  An observation is a point and an accepted biome set; it passes when the point's
 * biome is in the set (with MCB_SEARCH_TIES: when any least-distance leaf's biome is, the game's
 * tie nondeterminism).  Layout = McgObs (engine/cuda/mc_gpu.cuh:47): quart coordinates, or block
 * coordinates seen through BiomeManager when flags has MCB_OBS_BLOCK; flags' low 8 bits = the End
 * cell-window radius (12 = the game's window; smaller only where the set's answer cannot depend on
 * farther cells, see mcb_end_obs_radius).
 * ================================================================================================ */
typedef struct McbObs {
    int32_t x, y, z, flags;
    uint64_t mask[2];        /* bit b = biome id b accepted */
} McbObs;
#define MCB_OBS_BLOCK 0x100
#define MCB_MAX_OBS 64
/* Candidates: kind 0 = list[0..n), 1 = structural seeds base | i << 48 for i < 65536 (n ignored),
 * 2 = range base + i for i < n (McgSeedGen, engine/cuda/mc_gpu.cuh:695-702). */
typedef struct McbSeedSource {
    int32_t kind, pad;
    const int64_t *list;     /* host pointer */
    uint64_t base, n;
} McbSeedSource;
#define MCB_SEARCH_TIES 1    /* accept any least-distance leaf's biome */
#define MCB_SEARCH_NOPRUNE 2 /* Overworld: no early rejection on partial climate targets */
#define MCB_SEARCH_SLOTS 4   /* evaluate from full per-seed state (any dimension; the Nether and End always do) */
/* Observations are checked in the given order; a candidate stops at its first failure.  reached[j]
 * (nobs+1 entries) = candidates that passed exactly j observations; found = the candidates that
 * passed all, sorted as unsigned, at most max_found of them (*n_found; reached[nobs] counts all);
 * kernel_ms = device time.  Same result as gb_search over the same candidates. */
int mcb_search(McbEngine *e, const McbSeedSource *src, const McbObs *obs, int nobs, int flags, int64_t *found,
               uint32_t max_found, uint32_t *n_found, uint64_t *reached, double *kernel_ms);
/* End window radius for an accepted set (the smallest cell window that decides membership,
 * engine/cuda/mc_gpu.cuh:652-666); 0 when the answer does not depend on the island field. */
int mcb_end_obs_radius(const McbEngine *e, const uint64_t mask[2]);

/* Seeds x points, host buffers, synchronous (mcg_ref_points contract). */
int mcb_points(McbEngine *e, const int64_t *seeds, int ns, const int32_t *pts, int np, uint8_t *out_biome,
               int32_t *out_tg);
/* Same with device buffers, asynchronous on `stream` (a hipStream_t, NULL = default stream). */
int mcb_points_dev(McbEngine *e, const int64_t *d_seeds, int ns, const int32_t *d_pts, int np, uint8_t *d_out_biome,
                   int32_t *d_out_tg, void *stream);
/* mcb_points_dev that allocates nothing: no counterpart in the reference.  For callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * per-seed noise state and slot map live in one static device array per device (MCB_NOALLOC_MAX_SEEDS seed
 * slots, 15.8 MB, held until the process ends); ns > MCB_NOALLOC_MAX_SEEDS fails with MCB_E_LIMIT.  Same
 * results as mcb_points_dev.  Calls that overlap on one device share that array and race: order them
 * (one stream, or events).  The engine's device must be current. */
#define MCB_NOALLOC_MAX_SEEDS 1024
int mcb_points_dev_noalloc(McbEngine *e, const int64_t *d_seeds, int ns, const int32_t *d_pts, int np,
                           uint8_t *d_out_biome, int32_t *d_out_tg, void *stream);

/* A region request.  zoom = 0: quart grid, out[out_off + j*nx*nz + iz*nx + ix] = biome at quart
 * (x0+ix, y[j], z0+iz).  zoom = 1: block grid at block resolution through BiomeManager with the
 * SHA-256 hashed seed, biome at block (x0+ix, y[j], z0+iz) (engine/cuda/mc_gpu.cuh:96-118);
 * at most 4 layers for zoom, 8 otherwise. */
typedef struct McbRegion {
    int64_t seed;
    int32_t x0, z0, nx, nz;
    int32_t ny;
    int32_t y[8];
    int32_t zoom;
    int32_t world;      /* context world index (mcb_context_world); ignored by single-engine plans */
    int64_t out_off;
} McbRegion;

/* Bytes of output a request list needs (max over out_off + ny*nx*nz). */
int64_t mcb_regions_bytes(const McbRegion *reqs, int n);
/* Plain per-call path: one request, host output, synchronous (allocates, launches, copies back). */
int mcb_region(McbEngine *e, const McbRegion *req, uint8_t *out_host);

/* Batched plan over many requests of this engine's world: seed slots, tile list and launch shape are
 * fixed at creation; each run initialises the seeds' noise state and evaluates every tile in one
 * persistent grid launch (plus the zoom pass), asynchronously on `stream`, into device memory d_out.
 * tile_w: tile width in columns, optionally | cpl << 16 with cpl = columns evaluated per lane (1, 2, 4;
 *         each lane walks cpl columns along z, warm-starting the biome search from the previous one);
 *         tile = tile_w x (256/tile_w)*cpl columns; 0 = default 16, cpl 1.
 * blocks: persistent workgroups (0 = one per tile). */
int mcb_plan_create(McbEngine *e, const McbRegion *reqs, int n, int tile_w, int blocks, McbPlan **out);
int mcb_plan_run(McbPlan *p, uint8_t *d_out, void *stream);
/* One stage of a run: 0 = seed state, 1 = quart-grid tiles, 2 = quart layers of zoomed requests,
 * 3 = zoom pass.  mcb_plan_run = stages 0..3 in order on one stream. */
int mcb_plan_run_stage(McbPlan *p, int stage, uint8_t *d_out, void *stream);
/* Capture the run into a hipGraph once and replay it (same d_out each time). */
int mcb_plan_run_graph(McbPlan *p, uint8_t *d_out, void *stream);
int mcb_plan_info(const McbPlan *p, int64_t *tiles, int64_t *columns, int64_t *samples, int64_t *seeds);
/* Launch shape of world group `group` of a plan (0 for single-engine plans): tile width, columns per lane,
 * persistent workgroups (0 = one per tile), and whether its quart-grid launch runs the 26.3 kernel form
 * without PerlinNoise.wrap's range test (1: every octave input of its requests provably stays inside
 * +-2^24, where wrap is the identity; chosen at plan creation). */
int mcb_plan_shape(const McbPlan *p, int group, int *tile_w, int *cpl, int *blocks, int *wrap_free);
void mcb_plan_free(McbPlan *p);
/* Project-written autotuner: per world group of the plan, time the quart-grid launch for each tile
 * shape (tile_w in 8/16/32/64) and persistent-workgroup count (one per tile, or 2/4/8 per CU) and keep the
 * fastest.  Uses d_out as scratch output.  Writes the chosen tile_w of the first group to *tile_w. */
int mcb_plan_autotune(McbPlan *p, uint8_t *d_out, void *stream, int *tile_w, double *best_ms);

/* ================================================================================================
 * Contexts: several worlds (version x dimension x preset) on one device, for uneven mixed batches
 * and the serving loop.  A region's `world` field is an index returned by mcb_context_world.
 * ================================================================================================ */
typedef struct McbContext McbContext;
int mcb_context_create(McbContext **out, const char *data_dir);
void mcb_context_free(McbContext *c);
/* Index of the world (opened on first use), or a negative MCB_E_* code. */
int mcb_context_world(McbContext *c, int version, int dim, const char *preset);
/* Plan over regions of any worlds of the context: one group (seed init + persistent grid launch +
 * zoom pass) per world present, run back to back on one stream; mcb_plan_run/_graph/_autotune apply. */
int mcb_context_plan(McbContext *c, const McbRegion *reqs, int n, int tile_w, int blocks, McbPlan **out);

/* Engine of a context world (owned by the context). */
int mcb_context_world_engine(McbContext *c, int world, McbEngine **e);

/* ================================================================================================
 * Serving loop (continuous batching).  Requests of any context world are admitted in arrival order
 * into one of `lanes` in-flight batches while their output and zoom-staging bytes fit the lane's
 * device arena of `batch_bytes`; seed state is kept in a per-world device LRU of `slot_capacity`
 * slots (15,472 bytes each) so repeated seeds skip their initialisation; batches on different lanes
 * overlap on the device, and each lane has its own completion thread.  Results are copied into the
 * caller's host buffer (ny*nx*nz bytes, layout of McbRegion, out_off ignored).  Latency = submit to
 * result in the caller's buffer.  Defaults (argument <= 0): batch_bytes 8 MiB, lanes 4, 16384 slots.  No counterpart in the
 * reference (its closest is the terrain slot producer, libmcgen/src/gpu_bridge.c:605-760).
 * ================================================================================================ */
typedef struct McbServer McbServer;
int mcb_server_create(McbServer **out, McbContext *ctx, int64_t batch_bytes, int lanes, int slot_capacity);
int mcb_server_submit(McbServer *s, const McbRegion *req, uint8_t *host_out, int64_t *ticket);
int mcb_server_wait(McbServer *s, int64_t ticket, double *latency_ms);
int mcb_server_drain(McbServer *s);
int mcb_server_stats(McbServer *s, double *p50_ms, double *p99_ms, double *p999_ms, int64_t *completed,
                     double *slot_hit_rate, int64_t *batches);
/* The n slowest requests since the last reset, slowest first: out[5i..5i+4] = total latency, then its
 * phases: queued (submit -> admitted to a batch), launch (admitted -> the batch's work enqueued), device
 * (enqueued -> the batch's results in pinned memory), copy (-> in the caller's buffer), all in ms.  Returns
 * how many were written (the rest of out is zero). */
int mcb_server_slowest(McbServer *s, int n, double *out);
int mcb_server_reset_stats(McbServer *s);
void mcb_server_destroy(McbServer *s);

/* ================================================================================================
 * The engine's vanilla world for one seed in libmcgen GPU-plugin form (McgBiomeWorld of
 * include/mcgpu.h): the climate sampler written as a density-function program in libmcgen's lowering,
 * the seed's octave tables (computed by the device seed-state kernel), the biome tree breadth-first.
 * Feeds mcgpu_biome_new without libmcgen.  No counterpart in the reference: libmcgen builds these
 * descriptions from datapacks (libmcgen/src/gpu_bridge.c:246-315).  Free with mcb_plugin_world_free.
 * ================================================================================================ */
struct McgBiomeWorld;
int mcb_plugin_world(McbEngine *e, int64_t seed, struct McgBiomeWorld **out);
void mcb_plugin_world_free(struct McgBiomeWorld *w);
/* How a handle of mcgpu_biome_new evaluates its climate: 1 = the program is the vanilla Overworld climate
 * router (the graph mcb_plugin_world writes, in libmcgen's lowering) and runs through the engine's fused
 * column code, 0 = the density-function interpreter; the biomes are the same either way.  With the
 * environment variable MCB_PLUGIN_INTERPRET=1 set when the handle is made, every program is interpreted
 * (to compare the two paths).  No counterpart in the reference. */
int mcb_plugin_router(void *h);
/* The same recognition for a world description, on the host only (no device call): 1 when
 * mcgpu_biome_new would run it through the engine's column code. */
int mcb_plugin_router_world(const struct McgBiomeWorld *w);

/* A terrain world (McgTerrainWorld of include/mcgpu.h) over the engine's vanilla 26.3+ Overworld climate
 * for one seed, shaped like the game's Overworld final_density: depth from the y gradient and the offset
 * spline, a sloped body with a 3D noise, a cave noise, a top slide, interpolation over 4 x 8 x 4 cells and
 * a final squeeze.  The game's own graph is datapack JSON the engine does not carry, so this is a stand-in
 * of the same shape, not the vanilla function.  Feeds mcgpu_terrain_new without libmcgen (benchmarks,
 * tests).  No counterpart in the reference.  Free with mcb_plugin_terrain_world_free. */
struct McgTerrainWorld;
int mcb_plugin_terrain_world(McbEngine *e, int64_t seed, struct McgTerrainWorld **out);
void mcb_plugin_terrain_world_free(struct McgTerrainWorld *w);
/* mcgpu_terrain_batch with the density left in device memory for device-side consumers (no host copy,
 * no vein patches or cache cells): *dens = chunk i's 16 * nh * 16 floats at *dens + i * 16*nh*16, in
 * mcgpu_terrain_batch's layout, valid until the handle's next batch.  No counterpart in the reference. */
int mcb_terrain_batch_device(void *h, int n, const int *cx, const int *cz, const float **dens);

/* Whole chunks as libmcgen stores them (world_chunk_biomes, libmcgen/src/biome.c:125-157): for each of n
 * chunks (cx, cz) in cxz, qh*16 bytes out[c*qh*16 + (qy*4 + qz)*4 + qx], qy counted from the world's lowest
 * quart (Overworld: 96 layers from quart -16; Nether, End: 64 from 0).  Each chunk's fitness ties are
 * resolved along its fill order (sections bottom-up, then x, y, z) with the search history starting at
 * the chunk's first cell; the End per cell. */
int mcb_chunk_biomes(McbEngine *e, int64_t seed, int n, const int32_t *cxz, uint8_t *out);

#ifdef __cplusplus
}
#endif
#endif
