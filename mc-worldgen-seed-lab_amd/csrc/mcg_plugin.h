// mcg_plugin.h - internals of the libmcgen GPU plugin (include/mcgpu.h): the compiled form of a
// density-function program and of the biome R-tree, shared by the host side (mcg_plugin.cpp) and the
// kernels (mcg_plugin.hip).
//
// A McgProg arrives as a node list in evaluation order (children first).  The host compiles it once per
// world into PIns records whose operands are value slots, not node numbers: a linear scan over the
// nodes gives each value a slot from its definition to its last use, so a lane keeps only the values
// that are still live (a climate program of ~100 nodes needs ~10 slots) and the slots fit in LDS.
// Every lane of a wave evaluates the same instruction stream, so each record is read with scalar loads.
#ifndef MCG_PLUGIN_H
#define MCG_PLUGIN_H
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "../../include/mcgpu.h"
#include "mcb_common.h"

namespace mcb {
namespace plug {

constexpr int kWG = 256;                 // lanes per workgroup of the plugin kernels
constexpr int kChainCap = 12;            // tied cells followed back along the fill order (reference: 12)
constexpr int kSplineDepth = 16;         // spline nesting accepted (the reference's frame stack holds 16)
constexpr int kMaxLdsBytes = 64 * 1024;  // LDS per workgroup the kernels may use
// Staged permutation tables: single (256 B, masked indices) or doubled (512 B, no masks).  Single measured
// faster on MI355X (26.3 grid 3.43 -> 3.25 ms, 26.1/26.2 2048^2 step-4 grids 4.4 -> 3.8 ms): the smaller LDS
// footprint fits one more workgroup per CU, which outweighs the masks in this latency-bound kernel.
constexpr bool kPermDoubled = false;
// LDS gradient table of k_cells: 256 entries of 8 bytes (float noise) or 16 bytes (double noise)
constexpr int grad_lds_bytes(bool newf) { return newf ? 256 * 8 : 256 * 16; }

// one program node, compiled (128 bytes; operand fields are value slots)
// PIns.cm: bits 0-2 = the node's x/y/z override mask; kCmYFree = the node's value does not depend on
// the block y (skipped when a lane evaluates the same x and z again)
constexpr int32_t kCmYFree = 8;
struct PIns {
    int32_t op, dst, a, b, c;
    int32_t ns, sp, arr0, narr, thr0, nthr;
    int32_t i0, i1, i2, i3;
    int32_t cm, cx, cy, cz, oct;
    float f0, f1, f2, f3;
    double d0, d1;
    int32_t pad[4];
};
static_assert(sizeof(PIns) == 128, "PIns layout");

// R-tree node in pre-order with escape links (the reference's BFS array, walked depth-first in its child
// order).  d6 = squared distance of the seventh (offset) parameter to a target offset of 0.
struct PNode {
    int32_t lo[6], hi[6];
    uint64_t d6;
    int32_t escape, leaf, biome, pad;
};
static_assert(sizeof(PNode) == 72, "PNode layout");

// A leaf listed in a climate cell, in the compact form the cell table needs (its bounds fit 16 bits and its
// offset distance 32 when the table exists): the records follow the cell lists' order, one load each.
struct PRec {
    int16_t lo[6], hi[6];
    uint32_t d6;
    int32_t leaf;                  // the leaf's tree node | its biome << 16
};
static_assert(sizeof(PRec) == 32, "PRec layout");

// device view of one world
struct PWorld {
    const PIns *ins;
    const McgNoise *noise;
    const McgLayer *layers;
    const McgOct *octs;            // offsets (and the permutation tables when they are not staged)
    const McgSpline *sp;
    const float *splf;
    const int32_t *spc;
    const int32_t *arr;            // slot numbers (ISEL choices, spline coordinates)
    const double *thr;
    const PNode *tree;
    // climate cell table over the tree with strict lists (every leaf that can be nearest or tie in the
    // cell, in pre-order); centry null: every search walks the tree
    const int32_t *cbp;            // breakpoints of dimension k at cbp + coff[k], cnb[k] of them
    const uint32_t *centry;        // (first << 8) | count; count 0: walk the tree
    const uint16_t *clist;
    int32_t cnb[6], coff[6];
    // cell locator of targets within 16 bits (null: binary searches over cbp): per tree dimension 256
    // buckets of 256 values, entry = (cells below the bucket) * cstride | (offset of its one breakpoint) << 20
    const uint32_t *clut;
    int32_t cstride[6];
    const PRec *crec;              // the listed leaves, entry by entry
    // the entries again for the locator's path: count in bits 0-7; count 1: the one leaf that can be at
    // the least distance anywhere in the cell (tree node in bits 8-23, biome in bits 24-31); count > 1:
    // first record in crec (bits 8-31); count 0: walk the tree
    const uint32_t *cent1;
    int32_t nins, noct, ntree, nslot, spl_depth;
    int32_t nydep;                 // ins[nins .. nins + nydep): the y-dependent instructions again
    int32_t nnoise, nlayers;
    int32_t chain_cap;             // tied cells followed back before a cell is reported unresolved
    int32_t root_slot[6];          // -1: field absent
    int32_t source, fixed_biome, end_ids[5], min_qy, qh;
    int64_t zoom_seed;
    // the vanilla Overworld router (mcg_router.cpp): its constants in the engine's world form and its
    // octaves in the engine's seed layout; null: the program is interpreted
    const McbWorld *vw;
    const McbSlot *vslot;
    // End islands heights (end_height_j) of a lattice of chunks, filled by k_end_table before the launch:
    // chunk (etab_x0 + i % etab_nx, etab_z0 + i / etab_nx) at etab[i], for the islands nodes of octave
    // etab_oct; null: every islands node computes its window
    const float *etab;
    int32_t etab_oct, etab_x0, etab_z0, etab_nx, etab_nz;
};

// work description of one launch
enum { GEN_LATTICE = 0, GEN_POINTS = 1, GEN_GRIDPTS = 2, GEN_CHUNKS = 3, GEN_LIST = 4 };
struct PJob {
    int32_t gen;
    int64_t n;                     // items
    // lattice: cells (q0x + i % nqx, qys[layer], q0z + (i / nqx) % nqz), layer = i / (nqx * nqz); for the
    // End source a lattice of chunks (c0x + i % nqx, c0z + ..., qys[layer])
    int32_t q0x, q0z, nqx, nqz, qys[2];
    // grid points: block (x0 + ix * step, y, z0 + iz * step)
    int32_t x0, z0, nx, step, y;
    const int32_t *xyz;            // points: 3 ints per item; chunks: (cx, cz) per chunk
    int32_t cells;                 // chunks: quart cells per chunk (16 per quart layer), item = chunk * cells + cell
    uint8_t *out;
    int32_t *nbad;
    // list: the cells list[0 .. *nlist), each written to out[i]: lattice cells (i = layer * nqx * nqz +
    // column, as the lattice numbers them) when cq is null, else cell i at quart (cq[i], cq[nc + i],
    // cq[2 * nc + i]) with nc = ncells
    const int32_t *list;
    const int32_t *nlist;
    const int32_t *cq;
    int64_t ncells;
    int32_t geom;                  // router jobs: how cell i is placed (the job's own gen; GEN_LIST keeps it)
};

// Whole chunks of a world description in libmcgen's stored form (world_chunk_biomes,
// libmcgen/src/biome.c:125-157): out[c*cells + (qy*4 + qz)*4 + qx], each chunk's ties resolved along its
// fill order from the chunk's first cell with no length cap.  0 = ok, else err is set.
int chunk_biomes(const McgBiomeWorld *W, int n, const int32_t *cxz, uint8_t *out, int *nunres, char *err, size_t errlen);

// A compiled world kept between calls (the handle mcgpu_biome_new returns), with a chosen tie-chain cap:
// kChainCap for the plugin ABI, unbounded (1 << 30) for libmcgen's CPU semantics (biome.c:87-120 has none).
// The grid and points calls of the handle are mcgpu_biome_grid / mcgpu_biome_points.
void *world_new(const McgBiomeWorld *W, int chain_cap, char *err, size_t errlen);
int world_chunks(void *h, int n, const int32_t *cxz, uint8_t *out, int *nunres);

// Entry points of the plugin ABI run inside a PluginGate: the process-wide lock (calls are serialised, as
// the reference's are) and, when needs_device, the bound device made current (ready = false with the
// error set when it cannot be).  plugin_fail sets the thread's error text (mcgpu_last_error) and returns
// false.
struct PluginGate {
    std::unique_lock<std::mutex> lock;
    bool ready;
    explicit PluginGate(bool needs_device);
};
bool plugin_fail(const char *fmt, ...);

// The program of W is the vanilla Overworld climate router (mcg_router.cpp): w = its constants in the
// engine's world form, slot = its octaves in the engine's seed layout (permutations and offsets; the
// derived McbOct fields are completed on the device by finish_router_slot).  false: interpret the program.
bool vanilla_router(const McgBiomeWorld *W, McbWorld &w, McbSlot &slot);
hipError_t finish_router_slot(McbSlot *d_slot, hipStream_t st);

// launchers (mcg_plugin.hip); lds = bytes of dynamic LDS, perm_lds = permutation tables staged in LDS
hipError_t launch_cells(const PWorld &W, const PJob &J, bool newf, bool perm_lds, int lds, hipStream_t st);
// End islands heights of the chunks (x0 + i % nx, z0 + i / nx), i < nx * nz, from octave oct's simplex
// table, into out[i] (26.1/26.2 form when old)
hipError_t launch_end_table(const PWorld &W, int oct, int x0, int z0, int nx, int nz, bool old, float *out, hipStream_t st);
// Cells of a world with the vanilla router (W.vw) in launches: the quantized climate of every cell into
// tgt (planes of ncells int32: the six tree targets, then for points the cell's quart x, y, z); the table
// search of every cell at high occupancy, listing the cells it cannot settle (ties to follow along the fill
// order, targets beyond the table) in fix / cnt[0]; for the first kWinCells of those, every cell of their
// chain window (the cell and kWin - 1 cells back in fill order) evaluated in parallel into win, then each
// chain resolved from its window; chains the window cannot settle (uncapped ones that run past it) and
// listed cells beyond kWinCells through the per-lane state machine (list fix + ncells, length cnt[1]).
// J.gen: GEN_LATTICE (ncells = 2 * nqx * nqz; tgt holds 6 planes), GEN_CHUNKS (ncells = J.n, whole chunks;
// 6 planes), GEN_POINTS or GEN_GRIDPTS (ncells = J.n; 9 planes).  fix holds 2 * ncells, win kWinCells *
// kWin records; cnt[0..1] must be zero on entry.
constexpr int kWin = kChainCap + 1;
constexpr int kWinCells = 32768;
struct WinCell {
    int64_t tg[6];                 // the cell's targets
    uint64_t d;                    // its least distance (unless single)
    int32_t leaf, biome;           // first leaf at the least distance, its biome
    uint8_t ok, tied, single, pad; // ok: finite targets and a leaf found
    int32_t pad2;
};
static_assert(sizeof(WinCell) == 72, "WinCell layout");
int64_t van_cells(const PJob &J);
int van_planes(const PJob &J);
hipError_t launch_van_cells(const PWorld &W, const PJob &J, bool newf, int32_t *tgt, int32_t *fix, WinCell *win, int32_t *cnt,
                            hipStream_t st);
// block grid from a lattice of cells (MULTI) or of chunks (END): out[iz*nx+ix]
hipError_t launch_zoom_lattice(const PWorld &W, const PJob &L, int x0, int z0, int nx, int nz, int step, int y,
                               uint8_t *out, int32_t *nbad, hipStream_t st);

}  // namespace plug
}  // namespace mcb
#endif
