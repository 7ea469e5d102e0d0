/* mcoracle.h - CPU restatement of the reference biome-sampling path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  The product library never links or calls it.
 *
 * Every function restates the reference's algorithm from its source
 * (engine headers, engine/cuda/mcg_host.c, engine/cuda/mc_gpu.cuh,
 * libmcgen/src/region.c, libmcgen/src/biome.c); the file:line each one follows
 * is given beside its definition in mcoracle.c.  Build: oracle/Makefile
 * (gcc -O2 -ffp-contract=off -fopenmp, never -ffast-math).
 */
#ifndef MCORACLE_H
#define MCORACLE_H
#include <stdint.h>
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

enum { OR_V26_1 = 0, OR_V26_2 = 1, OR_V26_3 = 2, OR_V26_4 = 3 };   /* OR_V26_4 = 26.4-snapshot-2 */
enum { OR_OVERWORLD = 0, OR_NETHER = 1, OR_END = 2 };

typedef struct OrWorld OrWorld;

/* data_root holds climate-<V>.txt, params/<V>/{overworld,nether}.tsv.
 * preset: "normal" | "overworld" | "amplified" | "large_biomes" (Overworld only). */
OrWorld *or_open(const char *data_root, int version, int dim, const char *preset);
void or_close(OrWorld *w);
int or_mode(const OrWorld *w);                  /* 0 = double noise (26.1/26.2), 1 = float noise (26.3) */

/* 67 biome names in alphabetical order = the engine's u8 biome ids. */
int or_biome_count(void);
const char *or_biome_name(int id);
int or_biome_id(const char *name);

/* Same contract as mcg_ref_points: pts = (qx,qy,qz) triples; out_biome[s*np+p];
 * out_tg optional, 6 i32 per sample (t,h,c,e,d,w).  OpenMP over seeds. */
void or_points(const OrWorld *w, const int64_t *seeds, int ns, const int32_t *pts, int np,
               uint8_t *out_biome, int32_t *out_tg);

/* Quart grid out[iz*nx+ix] at (qx0+ix, qy, qz0+iz).
 * semantics 0: per-sample search with no candidate (engine).
 * semantics 1: row-major search history (each search starts from the previous
 *              sample's leaf, first sample from none).
 * semantics 2: the Java oracle's chunk grids: 26.3 climate through the chunk (volume) path, search
 *              history carried chunk by chunk (quart rows then columns inside a chunk).  */
void or_grid(const OrWorld *w, int64_t seed, int qx0, int qz0, int nx, int nz, int qy, int semantics,
             uint8_t *out);

/* Block grid with BiomeManager zoom: out[iz*nx+ix] = biome at block (x0+ix*step, y, z0+iz*step)
 * using engine search semantics at the zoomed quart cell. */
void or_block_grid(const OrWorld *w, int64_t seed, int x0, int z0, int nx, int nz, int step, int y,
                   uint8_t *out);

/* Climate values at a block position: out6 = T,H,C,E,D,W as float (what Climate.target quantizes). */
void or_climate_raw(const OrWorld *w, int64_t seed, int bx, int by, int bz, float out6[6]);
/* Same, with the pre-cast double values of the 26.1/26.2 path (float values widened for 26.3),
 * plus the offset density function: out7 = C,E,W(ridges),offset,depth,T,H. */
void or_climate_df(const OrWorld *w, int64_t seed, int bx, int by, int bz, double out7[7]);

/* One NormalNoise of the version's climate table by noise name ("temperature", "nether/vegetation" ...),
 * instantiated as RandomState does (positional Xoroshiro for overworld noises, legacy LCG for the
 * nether biome noises).  Returns the value as double (float widened for 26.3). */
int or_noise_value(const OrWorld *w, const char *noise_name, int64_t seed, double x, double y, double z,
                   double *out);
/* NormalNoise built from explicit parameters (26.3 form), positional instantiation: unit KAT helper. */
double or_noise_value_params(int version, const char *noise_name, int first_octave, int n_amp,
                             const double *amps, double base_amplitude, int64_t seed,
                             double x, double y, double z);

/* End: erosion density value at block (bx,bz) (what TheEndBiomeSource reads) and the biome id. */
double or_end_erosion(const OrWorld *w, int64_t seed, int bx, int bz);

/* BiomeManager: SHA-256 hashed seed and the zoomed quart cell of a block. */
int64_t or_hashed_seed(int64_t seed);
void or_zoom(int64_t hashed, int x, int y, int z, int32_t q[3]);

/* Tree inspection: number of nodes reachable from the root and a canonical
 * pre-order serialisation (for comparison with the game's tree dump):
 * per node: is_leaf, 14 bounds (lo,hi per dim), leaf param index or child count. */
int or_tree_serialize(const OrWorld *w, int64_t *buf, int cap);
/* Brute-force: all leaves at minimum distance; returns the count, biome mask in mask2[2]. */
int or_tie_mask(const OrWorld *w, const int32_t tg6[6], uint64_t mask2[2]);
/* Seed search over n candidates (see mcoracle.c); returns the number found */
int64_t or_search(const OrWorld *w, const int64_t *seeds, int64_t n, const int32_t *obs_xyz, const int32_t *obs_block,
                  const uint64_t *masks, int nobs, int ties, uint64_t *reached, int64_t *found, int64_t max_found);
/* End island field of one seed over a block grid: kind 0 height value at (bx/8, bz/8), 1 erosion */
void or_end_field(const OrWorld *w, int64_t seed, int bx0, int bz0, int nx, int nz, int step, int kind, double *out);
/* biome of quantized targets {t,h,c,e,d,w} by the tree search (engine/mc_biomes.h:31-62), n targets */
void or_targets(const OrWorld *w, const int32_t *tg6, int n, uint8_t *out);
/* Leaves visited per search (instrumentation): average over the given targets. */
double or_tree_visits(const OrWorld *w, const int32_t *tg6, int n);

/* splitmix64 stream used by the synthetic workloads: state += golden; return mix13(state). */
uint64_t or_splitmix64_next(uint64_t *state);

uint64_t or_fnv1a_names(const uint8_t *ids, size_t n);
/* whole chunks as libmcgen stores them (biome.c:125-157): per chunk qh*16 bytes [(qy*4+qz)*4+qx] */
void or_chunk_biomes(const OrWorld *w, int64_t seed, int n, const int32_t *cxz, uint8_t *out);

#ifdef __cplusplus
}
#endif
#endif
