/* mcgpu.h - libmcgen GPU plugin ABI (version 2), implemented for MI355X by libmcbiome_hip.so.
 *
 * libmcgen (the reference's world-generation library) loads an optional GPU library with dlopen and hands
 * it flat, plain-old-data descriptions of a world's biome source: a density-function program for the
 * six climate fields, the noise octave tables, the biome R-tree and the BiomeManager seed
 * (libmcgen/gpu/mcgen_gpu_abi.h:1-123; loader libmcgen/src/gpu_bridge.c:27-148).  This header declares
 * that ABI: the structures below must keep the reference's layout field for field (the loader passes
 * them by pointer), so their members, order and enum values are the reference's; everything behind them
 * is this library's own.  A libmcgen build picks this library up through MCGEN_CUDA_LIB (or
 * mcgen_gpu_set_library_path) with no change on its side.
 *
 * Semantics (what a caller can rely on):
 *  - mcgpu_biome_grid / mcgpu_biome_points return, for each position, the biome libmcgen's CPU path
 *    returns (mcgen_biome_at: BiomeManager zoom, climate of the chosen quart cell, R-tree search with the
 *    chunk-fill-order tie chain, libmcgen/src/biome.c:87-120), or 255 where it cannot be matched exactly
 *    (a non-finite or out-of-range climate value, a tie chain longer than 12 cells, an R-tree search
 *    that finds no leaf); *nunres counts the 255s and the caller recomputes those on the CPU.
 *  - Return codes: 0 = ok, 1 = error (text in mcgpu_last_error(), per thread).  *_new returns NULL and
 *    writes the reason into err.
 *  - Every entry point holds one process-wide lock (the reference serialises its entries the same way).
 */
#ifndef MCGPU_H
#define MCGPU_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define MCGPU_ABI 2

/* Node kinds of the 26.3+ (float) programs: the order of libmcgen's K_* (mcgen_gpu_abi.h:17-25). */
enum {
    MCGK_CONST, MCGK_NOISE, MCGK_NOISE_XZ, MCGK_NOISE_XYZ, MCGK_SHIFT_B, MCGK_END, MCGK_DIST, MCGK_GRAD_CLAMP,
    MCGK_GRAD_REPEAT, MCGK_GRAD_MIRROR, MCGK_CTX_ALPHA, MCGK_CTX_OFFSET, MCGK_CTX_BEARD, MCGK_ABS, MCGK_SQUARE,
    MCGK_CUBE, MCGK_SQRT, MCGK_LEAKY, MCGK_RECIP, MCGK_NEG, MCGK_SQUEEZE, MCGK_LOG, MCGK_SIGN, MCGK_ROUND_INT,
    MCGK_ROUND, MCGK_ADD, MCGK_CADD, MCGK_CSUB, MCGK_SUB, MCGK_MUL, MCGK_CMUL, MCGK_DIV, MCGK_CDIV, MCGK_MIN,
    MCGK_CMIN, MCGK_MAX, MCGK_CMAX, MCGK_POW_CB, MCGK_POW_CE, MCGK_POW, MCGK_SPLINE, MCGK_LERP, MCGK_LERP_CF,
    MCGK_LERP_CS, MCGK_CLAMP, MCGK_RANGE_C, MCGK_RANGE, MCGK_ISEL1, MCGK_ISEL, MCGK_CACHE, MCGK_BLEND_DENSITY,
    MCGK_INTERP, MCGK_SLICE_X, MCGK_SLICE_Y, MCGK_SLICE_Z, MCGK_SLICE_XZ, MCGK_FTS,
    MCGK__COUNT
};
/* Node kinds of the 26.1/26.2 (double) programs (mcgen_gpu_abi.h:27-33). */
enum {
    MCGO_CONST, MCGO_NOISE, MCGO_SHIFTED_NOISE, MCGO_SHIFT_A, MCGO_SHIFT_B, MCGO_SHIFT, MCGO_END_ISLANDS,
    MCGO_Y_GRADIENT, MCGO_WEIRD_SCALED, MCGO_BLEND_ALPHA, MCGO_ZERO, MCGO_ABS, MCGO_SQUARE, MCGO_CUBE,
    MCGO_HALF_NEG, MCGO_QUARTER_NEG, MCGO_RECIP, MCGO_SQUEEZE, MCGO_CLAMP, MCGO_ADD, MCGO_MUL, MCGO_MIN, MCGO_MAX,
    MCGO_CADD, MCGO_CMUL, MCGO_RANGE, MCGO_ISEL, MCGO_SPLINE, MCGO_PASS,
    MCGO__COUNT
};

/* One noise octave: coordinate offsets and permutation table (mcgen_gpu_abi.h:36). */
typedef struct McgOct { double xo, yo, zo; uint8_t p[256]; } McgOct;
/* One noise layer (mcgen_gpu_abi.h:40).  26.3+: a = frequency, amp_f = amplitude, kind 0 perlin / 1 smeared
 * with b = y fudge.  26.1/26.2 (one level of a legacy PerlinNoise): oct = -1 for an empty level,
 * a = input factor, b = value factor, amp_d = amplitude. */
typedef struct McgLayer { double a, b, amp_d; float amp_f; int oct; int kind; int pad; } McgLayer;
/* A noise: layers [l0, l0+n0) (26.3+: the whole stack; 26.1/26.2: the first PerlinNoise), [l1, l1+n1) the
 * second PerlinNoise and value_factor of a 26.1/26.2 NormalNoise (mcgen_gpu_abi.h:41-45). */
typedef struct McgNoise { int l0, n0; int l1, n1; double value_factor; } McgNoise;
/* Spline node: constant, or n points (splf[loc0..loc0+n) locations, splf[loc0+n..loc0+2n) derivatives,
 * spc[val0..val0+n) child splines) over coordinate `coord` of the owning node's list (mcgen_gpu_abi.h:48-54). */
typedef struct McgSpline { int is_const; float value; int coord, n; int loc0; int val0; } McgSpline;
/* Program node (mcgen_gpu_abi.h:56-68).  Children come before their parents in nodes[]. */
typedef struct McgNode {
    int k;
    int a, b, c;
    int ns;
    int sp;
    int arr0, narr;
    int thr0, nthr;
    int i0, i1, i2, i3;
    float f0, f1, f2, f3;
    double d0, d1;
    int cm, cx, cy, cz;   /* coordinates fixed by slices: mask (1 x, 2 y, 4 z) and values */
    int oct;              /* END: simplex octave */
    int cid;              /* CACHE: cache number */
} McgNode;

#define MCG_MAX_ROOTS 16
typedef struct McgProg {
    int old;                     /* 0: float program (26.3+), 1: double program (26.1/26.2) */
    int nnodes;  McgNode *nodes;
    int nnoise;  McgNoise *noise;
    int nlayers; McgLayer *layers;
    int noct;    McgOct *octs;
    int nsp;     McgSpline *sp;
    int nsplf;   float *splf;
    int nspc;    int *spc;
    int narr;    int *arr;
    int nthr;    double *thr;
    int nroots;  int root[MCG_MAX_ROOTS];
    int ncache;
} McgProg;

/* R-tree node: children [first, first+count), a leaf has count 0 (mcgen_gpu_abi.h:90). */
typedef struct McgTreeNode { int32_t lo[7], hi[7]; int32_t first, count, biome, pad; } McgTreeNode;

enum { MCG_BS_MULTI = 0, MCG_BS_END = 1, MCG_BS_FIXED = 2 };
/* World description for biome grids (mcgen_gpu_abi.h:93-103).  prog roots: 0 temperature,
 * 1 vegetation, 2 continents, 3 erosion, 4 depth, 5 ridges. */
typedef struct McgBiomeWorld {
    int source;
    int fixed_biome;
    int end_ids[5];              /* the_end, end_highlands, end_midlands, small_end_islands, end_barrens */
    int newf;                    /* 1: float program (26.3+) */
    int min_qy, qh;              /* quart y is clamped into [min_qy, min_qy + qh) */
    int64_t zoom_seed;           /* BiomeManager seed (obfuscated) */
    const McgProg *prog;
    int ntree; const McgTreeNode *tree; int tree_root;
} McgBiomeWorld;

/* World description for terrain (26.3+; mcgen_gpu_abi.h:105-123): the volume density program (root
 * root_density = final_density, vein roots by index into vol->root), the veins' point "gap" program, the
 * noise volume's y range [nmin, nmin + nh), the ore veins' positional random ("minecraft:ore"), the cache
 * numbers whose cell values the CPU needs after the density (aquifer), and the function whose volume the
 * CPU evaluates before the density (pre_root, -1 = none) with that volume's shape. */
#define MCG_MAX_VEINS 4
typedef struct McgVein { int density_root, richness_root, gap_root; int ore, raw, filler; float raw_chance; } McgVein;
typedef struct McgTerrainWorld {
    const McgProg *vol;
    const McgProg *gap;
    int nmin, nh;
    int root_density;
    int nveins; McgVein vein[MCG_MAX_VEINS];
    uint64_t ore_lo, ore_hi;
    int veins_on;
    int cache_ids[32]; int ncache_ids;
    int pre_root;
    int pre_sx, pre_sy, pre_sz, pre_dx, pre_dy, pre_dz, pre_y0, pre_xoff, pre_zoff;
} McgTerrainWorld;
/* Beardifier callback: fills out with the volume vol = {sx, sy, sz, x0, y0, z0, dx, dy, dz} (block
 * coordinates) for chunk chunk_index of the batch; returns 0 when the chunk has no Beardifier. */
typedef int (*McgBeardFn)(void *ud, int chunk_index, const int vol[9], float *out);

/* ---- exports bound by libmcgen/src/gpu_bridge.c:123-140 (all required) ---- */
int mcgpu_abi(void);                                                       /* MCGPU_ABI */
int mcgpu_device_count(void);
/* name, architecture major/minor (gfx950 reports 9 / 5), memory in MiB; 0 = ok */
int mcgpu_device_info(int index, char *name, size_t namelen, int *cc_major, int *cc_minor, size_t *mem_mb);
int mcgpu_select(int device);                                              /* once, before any work */
int mcgpu_selected(void);                                                  /* -1 before mcgpu_select */
const char *mcgpu_last_error(void);
int mcgpu_mem_info(size_t *free_mb, size_t *total_mb);
/* Copies everything it needs from *world (the caller may free it on return). */
void *mcgpu_biome_new(const McgBiomeWorld *world, char *err, size_t errlen);
void mcgpu_biome_free(void *h);
/* out[iz*nx+ix] = biome at block (x0 + ix*step, y, z0 + iz*step) (mcgpu_biome.cu:203-249). */
int mcgpu_biome_grid(void *h, int x0, int z0, int nx, int nz, int step, int y, unsigned char *out, int *nunres);
/* out[i] = biome at block (xyz[3i], xyz[3i+1], xyz[3i+2]) (mcgpu_biome.cu:252-280). */
int mcgpu_biome_points(void *h, int n, const int *xyz, unsigned char *out, int *nunres);

/* ---- terrain exports, optional for libmcgen (bound by gpu_bridge.c:141-144, used by gpu_bridge.c:503-656) ----
 * The volume executor (mcgpu_terrain.cu:556-862): mcgpu_terrain_new plans the world's final_density (and
 * vein roots) once; max_chunks caps the chunks of one batch (the result may be smaller: device memory).
 * Node kinds log, pow, round and find_top_surface, and interpolation or slices over volume shapes other
 * than the terrain's own, are refused at creation (the reference refuses the same).  NULL + err on
 * failure.  Copies everything it needs from *world. */
void *mcgpu_terrain_new(const McgTerrainWorld *world, int max_chunks, char *err, size_t errlen);
void mcgpu_terrain_free(void *h);
int mcgpu_terrain_max_chunks(void *h);
/* floats of cache-cell values per chunk in out_cells, and cache k's (in world->cache_ids order) number,
 * candidate volumes, their largest size and its offset in a chunk's cell floats; 0 = ok, -1 = no such k */
size_t mcgpu_terrain_cell_floats(void *h);
int mcgpu_terrain_cell_info(void *h, int k, int *cid, int *ncand, size_t *cap, size_t *off);
/* candidate j of cache k: vol = {sx, sy, sz, x, y0, z, dx, dy, dz}, x (z) relative to the chunk's block
 * origin when *xrel (*zrel); *n = its size */
int mcgpu_terrain_cell_cand(void *h, int k, int j, int vol[9], int *xrel, int *zrel, int *n);
/* pinned host memory for batch outputs (hipHostMalloc, portable) */
void *mcgpu_host_alloc(size_t bytes);
void mcgpu_host_free(void *p);
/* n <= max_chunks chunks (cx[i], cz[i]): out_dens = n x (16 * nh * 16) floats, chunk i's block (x, y, z) at
 * i * 16*nh*16 + (y - nmin) + (x + z*16) * nh; out_veins (or NULL) = n x (nh * 256) patches at
 * ((y - nmin)*16 + z)*16 + x, 0xFFFF = none; out_cells / out_which (or NULL) = per chunk the cache cells'
 * values and, per cache, which candidate volume the CPU would hold after the density (-1: the cell is
 * unset or as the pre-density volume left it, -2: unknown, sample the chunk on the CPU).  beard_cb (or
 * NULL) fills Beardifier volumes.  0 = ok, 1 = bad arguments, 3 = device error (mcgpu_last_error). */
int mcgpu_terrain_batch(void *h, int n, const int *cx, const int *cz, McgBeardFn beard_cb, void *beard_ud, float *out_dens,
                        unsigned short *out_veins, float *out_cells, signed char *out_which);

#ifdef __cplusplus
}
#endif
#endif
