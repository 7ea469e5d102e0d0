/* mcgplugin.h - CPU restatement of libmcgen's biome lookup on a McgBiomeWorld (test infrastructure;
 * see mcgplugin.c).  The world builder or_mcg_world lives in mcoracle.c. */
#ifndef MCGPLUGIN_H
#define MCGPLUGIN_H
#include <stdint.h>

#include "../include/mcgpu.h"
#include "mcoracle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* the vanilla world of `w` for one seed in plugin form (program, octaves, tree); free with or_mcg_world_free */
McgBiomeWorld *or_mcg_world(const OrWorld *w, int64_t seed);
void or_mcg_world_free(McgBiomeWorld *W);

/* the six root values at a block (float programs widened to double) */
int or_mcg_climate(const McgBiomeWorld *W, int bx, int by, int bz, double out6[6]);
/* biome of a quart cell as libmcgen's world_biome_cell; *unres = 1 where the plugin returns 255,
 * *chain = number of tied cells followed back */
int or_mcg_cell(const McgBiomeWorld *W, int qx, int qy, int qz, int *unres, int *chain);
/* one chunk's cells searched sequentially in fill order, each search starting from the previous leaf:
 * out[(qy*4 + qz)*4 + qx] (qy relative to min_qy), valid = climate accepted by the plugin rules */
int or_mcg_chunk_seq(const McgBiomeWorld *W, int cx, int cz, int32_t *out, int32_t *valid);
/* quart grid (qx0+ix, qy, qz0+iz) in row-major order with search history (Java oracle grids) */
int or_mcg_quart_history(const McgBiomeWorld *W, int qx0, int qz0, int nx, int nz, int qy, uint8_t *out);
/* mcgpu_biome_points / mcgpu_biome_grid semantics (255 = unresolved, counted in *nunres); cap = longest tie
 * chain followed (12: the plugin ABI; -1: none, libmcgen's CPU mcgen_biome_at) */
int or_mcg_points(const McgBiomeWorld *W, int n, const int *xyz, uint8_t *out, int *nunres, int cap);
int or_mcg_grid(const McgBiomeWorld *W, int x0, int z0, int nx, int nz, int step, int y, uint8_t *out, int *nunres, int cap);
/* the value of node `node` of a 26.3+ program at a block (point path, no caches), and a NoiseStack value
 * (noise.c:339-347) at a position: the point pieces of the volume restatement in mcgterrain.c */
float or_mcg_point_f(const McgProg *P, int node, int x, int y, int z);
float or_mcg_stack_f(const McgProg *P, int ns, double x, double y, double z);

#ifdef __cplusplus
}
#endif
#endif
