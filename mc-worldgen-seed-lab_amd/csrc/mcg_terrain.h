// mcg_terrain.h - internals of the terrain volume executor behind mcgpu_terrain_* (include/mcgpu.h):
// the plan the host compiles from a McgTerrainWorld (mcg_terrain.cpp) and the kernels that run it
// (mcg_terrain.hip).
//
// A chunk's final_density is a DAG of (node, volume) values: sampleVolume (libmcgen/src/df_new.c:917-1150)
// evaluates each node over a volume, and interpolation and slices ask their input for a different volume
// (the cell corners; the volume with one axis fixed).  The host walks that recursion once per world,
// relative to the chunk origin, and groups the values into passes: one pass per (volume, level), where a
// value's level is one more than that of any value it reads from another volume.  A pass is one launch,
// a lane per (chunk, element) of its volume, running the pass's instruction list with live values in
// LDS slots (slot s of lane l at s * 256 + l).  Only values read by a later pass or returned to the caller
// are written to HBM (the batch arena); everything else stays in LDS.  The reference runs one
// element-wise kernel per node and writes every node's whole volume to memory
// (libmcgen/gpu/mcgpu_terrain.cu:38-328, 414-550).
#ifndef MCG_TERRAIN_H
#define MCG_TERRAIN_H
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mcg_plugin.h"

namespace mcb {
namespace plug {

constexpr int kTWG = 256;                   // lanes per workgroup of the terrain kernels

// A volume relative to a chunk: x (z) starts at the chunk's block origin + xoff (zoff) when xrel (zrel),
// else at xoff (zoff); element (ix, iy, iz) is at index iy + (ix + iz * sx) * sy.
struct TVol {
    int32_t sx, sy, sz, dx, dy, dz, y0, xrel, xoff, zrel, zoff;
    bool operator==(const TVol &o) const {
        return sx == o.sx && sy == o.sy && sz == o.sz && dx == o.dx && dy == o.dy && dz == o.dz && y0 == o.y0 &&
               xrel == o.xrel && xoff == o.xoff && zrel == o.zrel && zoff == o.zoff;
    }
    __host__ __device__ int32_t size() const { return sx * sy * sz; }
};

// instruction kinds beyond the node kinds (MCGK_*): reads of values another pass wrote
enum { TOP_LOAD = 100, TOP_INTERP = 101, TOP_SLICE = 102 };

// One instruction of a pass.  a, b, c (and the slot list at arr0 for splines and interval_select) are
// LDS slots; store / src are arena offsets in floats of the value's batch buffer (chunk c, element e at
// offset + c * per + e, per = the volume size of that value), -1 when none.
struct TIns {
    int32_t op, dst, a, b, c;
    int32_t ns, sp, arr0, narr, thr0, nthr;
    int32_t i0, i1, i2, i3, oct;
    int32_t flag;                           // spline instance of the used-coordinate masks, -1: none
    int32_t g0, g1, g2, g3, g4;             // fetch geometry: see the kernel
    float f0, f1, f2, f3;
    int64_t store, src;
    double d0, d1;
};
static_assert(sizeof(TIns) % 8 == 0, "TIns alignment");

// the program's tables on the device
struct TProg {
    const McgNoise *noise;
    const McgLayer *layers;
    const McgOct *octs;
    const McgSpline *sp;
    const float *splf;
    const int32_t *spc;
    const double *thr;
};

// node kinds a pass holds, beyond arithmetic, fetches and coordinates: noises (and End islands), splines
enum { TF_NOISE = 1, TF_SPLINE = 2 };
constexpr int kShallowSpline = 6;           // spline nesting of the lean spline kernel

// one pass of one batch
struct TPass {
    const TIns *ins;
    const int32_t *arr;                     // slot lists
    int32_t nins, nslot;
    TVol v;
    int32_t n;                              // elements per chunk
    int32_t feat, spd;                      // TF_* of its instructions, deepest spline nesting
};
struct TBatch {
    const int32_t *cx0, *cz0;               // block origin of each chunk
    int32_t nchunks;
    float *arena;
    uint32_t *flags;                        // [spline instance][chunk] masks (null: not collected)
    int32_t flag_stride;                    // chunks per spline instance in flags
};

// ore-vein rule evaluation over the chunk volume (16 x nh x 16 from nmin), out[c * nd + (y*16 + z)*16 + x]
struct TVeins {
    int32_t n;                              // veins
    int64_t dens[MCG_MAX_VEINS], rich[MCG_MAX_VEINS];   // arena offsets (per chunk stride nd)
    int32_t ore[MCG_MAX_VEINS], raw[MCG_MAX_VEINS], filler[MCG_MAX_VEINS], gap_slot[MCG_MAX_VEINS];
    float raw_chance[MCG_MAX_VEINS];
    uint64_t lo, hi;                        // ore PositionalRandomFactory
    int32_t nmin, nh;
};

// item of the cache-cell gather: n floats from the arena at src to the staging buffer at dst
struct TGather {
    int64_t src, dst;
    int32_t n, pad;
};

hipError_t launch_tpass(const TProg &P, const TPass &A, const TBatch &B, hipStream_t st);

// A pass that is one interpolation over a cell-aligned unit-step volume followed by element-wise unary
// nodes (the usual tail of a final_density): each lane takes one run of cy blocks up a cell column, reads
// the cell's eight corners once and walks the run with the CPU's running y sum (df_new.c:867-888),
// applying the tail to each block.  The values equal the general pass's, operation for operation.
constexpr int kMaxTail = 8;
struct TInterp {
    int64_t src, store;                     // corner buffer, output buffer (arena offsets)
    int32_t ncv, cvx, cvy, cvz;             // corner floats per chunk, corner volume size
    int32_t lx, ly;                         // log2 of the cell size (x and z, y)
    float f0, f1;                           // 1 / cell size (x and z, y)
    int32_t ntail;
    int32_t op[kMaxTail];
    float k0[kMaxTail], k1[kMaxTail], k2[kMaxTail], k3[kMaxTail];
};
hipError_t launch_tinterp(const TInterp &I, const TVol &v, const TBatch &B, hipStream_t st);
// gap program: a point world (PWorld with the program's instructions) evaluated per block where needed
hipError_t launch_tveins(const PWorld &G, int lds, const TVeins &V, const TBatch &B, uint16_t *out, hipStream_t st);
hipError_t launch_tgather(const TGather *items, int nitems, const float *arena, float *dst, hipStream_t st);

// point program of a 26.3+ McgProg on the device (the ore-vein gap functions): compiled as the biome
// worlds' programs are, uploaded with one copy into *d_data (kept when its *cap bytes suffice, else
// allocated again; freed by the caller).  lds = dynamic LDS bytes of an evaluation (gradient table and
// value slots).  false with the plugin error set on failure.
bool point_program(const McgProg *p, PWorld &W, char **d_data, size_t *cap, int *lds);

}  // namespace plug
}  // namespace mcb
#endif
