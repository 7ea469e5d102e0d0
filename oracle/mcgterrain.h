/* mcgterrain.h - CPU restatement of libmcgen's 26.3+ terrain volume path on an exported program (test
 * infrastructure; see mcgterrain.c). */
#ifndef MCGTERRAIN_H
#define MCGTERRAIN_H
#include <stdint.h>

#include "../include/mcgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* sampleVolume of program node `node` over vol = {sx, sy, sz, x0, y0, z0, dx, dy, dz} in a fresh sampler
 * context (caches on).  0 = ok, -1 = a node kind or volume geometry the restatement does not cover. */
int or_terrain_volume(const McgProg *P, int node, const int vol[9], float *out);

/* One chunk (cx, cz) of a terrain world as the CPU terrain path samples it and as mcgpu_terrain_batch
 * reports it: out_dens = final_density over 16 x nh x 16 blocks from nmin (layout iy + (ix + iz*16)*nh);
 * out_veins (may be NULL) = the ore-vein patch per block ((y*16 + z)*16 + x, 0xFFFF = none, computed when
 * T->veins_on); for each T->cache_ids[k]: set[k] / changed[k] / cvol[9k..] / the values at cvals + k*cap,
 * as the CPU's sampler context holds them after the density (changed: not the state the pre volume left).
 * beard / ud: the Beardifier callback of the batch (may be NULL), called with chunk_index.  0 = ok. */
int or_terrain_chunk(const McgTerrainWorld *T, int cx, int cz, int chunk_index, McgBeardFn beard, void *ud,
                     float *out_dens, uint16_t *out_veins, int *set, int *changed, int *cvol, float *cvals, int cap);

/* Mth.getSeed(x, y, z) positional Xoroshiro of (lo, hi) and its first k floats */
void or_terrain_ore_floats(uint64_t lo, uint64_t hi, int x, int y, int z, int k, float *out);

#ifdef __cplusplus
}
#endif
#endif
