/* mcgen.h - the biome part of libmcgen's public C ABI (libmcgen/include/mcgen.h:1-172), exported by
 * libmcbiome_hip.so with libmcgen's names, signatures, return codes and layouts, for callers that use
 * libmcgen for biomes (map previews, chunk biome arrays).
 *
 * What is the same: every function below has libmcgen's signature; biomes follow libmcgen's CPU semantics
 * (mcgen_biome_at: BiomeManager zoom, quart y clamped into the world, fitness ties resolved along the
 * chunk's fill order with no cap, libmcgen/src/biome.c:87-120; region biomes: world_chunk_biomes,
 * biome.c:125-157), computed on the GPU for the vanilla worlds of 26.1, 26.2 and 26.3.  Biome ids are
 * libmcgen's: the version's biome registry sorted by name (libmcgen/src/pack.c:134-146), i.e. the biomes
 * the version's Overworld and Nether place, the five End biomes and the_void (26.1: 65, 26.2: 66, 26.3: 67
 * ids); mcgen_biome_count / mcgen_biome_name describe the registry of the McGen's version.
 *
 * What differs (this library implements libmcgen's biome stage only):
 *  - pack_dir is not read: the vanilla climate tables built into the library are used (version "26.1",
 *    "26.2" or "26.3"; "26.4-snapshot-2" and others -> MCGEN_E_VERSION);
 *  - world tweaks are refused (ntweaks > 0 -> MCGEN_E_UNSUPPORTED), mcgen_tweak_count is 0;
 *  - mcgen_generate_region computes MC_STAGE_BIOMES only (other stages -> MCGEN_E_UNSUPPORTED), so
 *    mcgen_region_blocks / mcgen_region_heightmap return NULL; there are no block states or structures;
 *  - everything runs on the GPU: mcgen_biome_grid_cpu and mcgen_gpu_set_compute(MCGEN_COMPUTE_CPU) return
 *    MCGEN_E_UNSUPPORTED, mcgen_gpu_biome_points reports *n_cpu = 0.
 */
#ifndef MCGEN_H
#define MCGEN_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define MCGEN_ABI_VERSION 1

typedef struct McGen McGen;
typedef struct McWorld McWorld;
typedef struct McRegion McRegion;

enum {
    MCGEN_OK = 0, MCGEN_E_ARG = 1, MCGEN_E_IO = 2, MCGEN_E_DATA = 3, MCGEN_E_VERSION = 4,
    MCGEN_E_NOMEM = 5, MCGEN_E_CANCEL = 6, MCGEN_E_UNSUPPORTED = 7, MCGEN_E_INTERNAL = 8
};
enum {
    MC_STAGE_BIOMES = 1u, MC_STAGE_TERRAIN = 2u, MC_STAGE_SURFACE = 4u, MC_STAGE_CARVERS = 8u,
    MC_STAGE_FEATURES = 16u, MC_STAGE_STRUCTURES = 32u, MC_STAGE_ALL = 63u
};

typedef struct McSeeds { int64_t climate, terrain, structures, features; } McSeeds;
static inline McSeeds mcgen_seeds_unified(int64_t s) { McSeeds r = { s, s, s, s }; return r; }
typedef struct McTweakInfo {
    const char *id, *label, *group, *description;
    double def, min, max, soft_min, soft_max;
    int is_int;
} McTweakInfo;
typedef struct McTweakValue { const char *id; double value; } McTweakValue;
typedef int (*McProgressFn)(void *ud, double fraction, const char *what);

/* ---- version, data (mcgen.h:64-86) ---- */
const char *mcgen_version(void);
int mcgen_abi_version(void);
int mcgen_open(const char *pack_dir, const char *version, McGen **out, char *err, size_t errlen);
void mcgen_close(McGen *g);
int mcgen_dimension_count(const McGen *g);
const char *mcgen_dimension_name(const McGen *g, int i);
int mcgen_preset_count(const McGen *g, const char *dimension);
const char *mcgen_preset_name(const McGen *g, const char *dimension, int i);
int mcgen_block_state_count(const McGen *g);
const char *mcgen_block_state_name(const McGen *g, int id);
int mcgen_block_state_from_name(const McGen *g, const char *name);
int mcgen_biome_count(const McGen *g);
const char *mcgen_biome_name(const McGen *g, int id);
int mcgen_tweak_count(const McGen *g);
const McTweakInfo *mcgen_tweak_info(const McGen *g, int i);

/* ---- world (mcgen.h:88-104) ---- */
int mcgen_world_new(McGen *g, const char *dimension, const char *preset, const McSeeds *seeds,
                    const McTweakValue *tweaks, int ntweaks, McWorld **out, char *err, size_t errlen);
void mcgen_world_free(McWorld *w);
int mcgen_world_set_schedule(McWorld *w, const char *path, char *err, size_t errlen);
int mcgen_world_min_y(const McWorld *w);
int mcgen_world_height(const McWorld *w);
int mcgen_world_sea_level(const McWorld *w);
int mcgen_biome_at(const McWorld *w, int x, int y, int z);        /* -1 on a NULL world or an error */
int mcgen_biome_grid(const McWorld *w, int x0, int z0, int nx, int nz, int step, int y, uint8_t *out);

/* ---- region (mcgen.h:106-120) ---- */
typedef struct McRegionInfo { int cx0, cz0, nx, nz, min_y, height; } McRegionInfo;
enum { MC_HM_WORLD_SURFACE = 0, MC_HM_OCEAN_FLOOR = 1, MC_HM_MOTION_BLOCKING = 2, MC_HM_MOTION_BLOCKING_NO_LEAVES = 3 };
int mcgen_generate_region(McWorld *w, int cx0, int cz0, int nx, int nz, uint32_t stages, int threads,
                          McProgressFn cb, void *ud, McRegion **out, char *err, size_t errlen);
void mcgen_region_free(McRegion *r);
void mcgen_region_info(const McRegion *r, McRegionInfo *info);
uint16_t *mcgen_region_blocks(McRegion *r, int cx, int cz);
uint8_t *mcgen_region_biomes(McRegion *r, int cx, int cz);       /* (height/4)*16 bytes, [qy][qz][qx] */
int16_t *mcgen_region_heightmap(McRegion *r, int cx, int cz, int kind);

/* ---- GPU controls (mcgen.h:146-172) ---- */
enum { MCGEN_COMPUTE_CPU = 0, MCGEN_COMPUTE_GPU = 1, MCGEN_COMPUTE_AUTO = 2 };
enum { MCGEN_GPU_BIOMES = 1, MCGEN_GPU_TERRAIN = 2 };
int mcgen_gpu_set_features(int mask);
int mcgen_gpu_get_features(void);
int mcgen_gpu_set_library_path(const char *path);
int mcgen_gpu_device_count(void);
int mcgen_gpu_device_info(int index, char *name, size_t namelen, int *cc_major, int *cc_minor, size_t *mem_mb);
int mcgen_gpu_set_compute(int mode, int device);
int mcgen_gpu_get_compute(void);
int mcgen_gpu_status(char *buf, size_t buflen);
int mcgen_gpu_selftest(McGen *g, int npoints, char *report, size_t replen);
int mcgen_gpu_biome_grid(const McWorld *w, int x0, int z0, int nx, int nz, int step, int y, uint8_t *out);
int mcgen_biome_grid_cpu(const McWorld *w, int x0, int z0, int nx, int nz, int step, int y, uint8_t *out);
int mcgen_gpu_biome_points(const McWorld *w, int n, const int *xyz, uint8_t *out, int *n_cpu);

#ifdef __cplusplus
}
#endif
#endif
