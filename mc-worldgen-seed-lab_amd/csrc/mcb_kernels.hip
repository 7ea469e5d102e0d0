// mcb_kernels.hip - HIP kernels of the MI355X biome engine (gfx950, wave64).
//
// Precision follows the reference CPU engine exactly (engine/mc_noise.h, engine/mc_climate.h):
// 26.1/26.2 noise in double, 26.3 noise with double coordinates and float arithmetic.  The build
// passes -ffp-contract=off and no fast-math flags, so every multiply and add below is a separate
// IEEE operation as in the reference; divisions and square roots are the correctly rounded ones.
//
// Kernels
//   k_seed_ow / k_seed_ne / k_seed_end : per-seed noise state (Xoroshiro or legacy LCG octave init,
//                                       Fisher-Yates permutations built in a per-lane LDS tile)
//   k_grid<MODE,DIM,MAXL>              : quart-grid tiles of many requests; the seed's 46 permutation
//                                       tables are staged in LDS, one column per lane, the climate of a
//                                       column is computed once and shared by all its y layers
//   k_grid_end<MODE>                   : End quart-grid tiles over an LDS list of the tile's island cells
//   k_points<MODE,DIM>                 : seeds x arbitrary quart points (the mcg_ref_points contract)
//   k_targets<DIM>                     : biomes of quantized climate targets (the search step alone)
//   k_end_field<MODE>                  : End island height / erosion field over block grids of many seeds
//   k_zoom                             : BiomeManager block zoom over quart layers written by k_grid
#include "mcb_device.h"

namespace mcb {
namespace {
using namespace dev;

// ------------------------------------------------------------------------------------------------
// Overworld seed state: one lane per (seed, octave).  RandomState: root = Xoroshiro(seed).forkPositional();
// noise = root.fromHashOf("minecraft:<noise>"); each stack = noise.forkPositional(); octave =
// stack.fromHashOf("octave_<n>")  (engine/mc_climate.h:97-106, engine/mc_noise.h:257-265).
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_seed_ow(const McbWorld *__restrict__ W, const int64_t *__restrict__ seeds,
                                               const int32_t *__restrict__ slot_of, int n, McbSlot *__restrict__ slots) {
    __shared__ uint32_t tile[64 * 64];
    const int lane = threadIdx.x;
    const int n_oct = W->n_oct;
    const int gid = blockIdx.x * 64 + lane;
    if (gid >= n * n_oct) return;                        // no barriers below: early exit is safe
    const int si = gid / n_oct, o = gid - si * n_oct;
    int nn = 0;
    while (nn + 1 < W->n_noise && o >= W->oct_base[nn + 1]) nn++;
    const McbNoise &N = W->noise[nn];
    const int rel = o - W->oct_base[nn], st = rel / N.nlev, k = rel - st * N.nlev;
    const int64_t seed = seeds[si];
    Xo root = xo_from_seed(seed);
    Xo pos = xo_fork(root);
    Xo r = xo_make(W->nlo[nn] ^ pos.a, W->nhi[nn] ^ pos.b);
    Xo s0 = xo_fork(r);
    Xo s1 = xo_fork(r);
    Xo sp = st ? s1 : s0;
    Xo rk = xo_make(N.hlo[k] ^ sp.a, N.hhi[k] ^ sp.b);
    double xo = xo_double(rk) * 256.0, yo = xo_double(rk) * 256.0, zo = xo_double(rk) * 256.0;
    fisher_yates(tile, lane, [&](int b) { return xo_bound(rk, (uint32_t)b); });
    McbSlot *slot = slots + slot_of[si];
    store_octave(slot, o, tile, lane, xo, yo, zo);
    if (o == 0) slot->seed = seed;
}

// ------------------------------------------------------------------------------------------------
// Nether seed state: LegacyFbm from java.util.Random(seed + noise): per stack the zero-index octave is
// drawn first (kept only if present), then indices zero-1 .. 0, absent ones skip 262 draws
// (engine/mc_noise.h:267-282; skip by LCG jump).  One lane per (seed, noise).
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_seed_ne(const McbWorld *__restrict__ W, const int64_t *__restrict__ seeds,
                                               const int32_t *__restrict__ slot_of, int n, McbSlot *__restrict__ slots) {
    __shared__ uint32_t tile[64 * 64];
    const int lane = threadIdx.x;
    const int gid = blockIdx.x * 64 + lane;
    if (gid >= n * 2) return;
    const int si = gid >> 1, nn = gid & 1;
    const McbNoise &N = W->noise[nn];
    const int64_t seed = seeds[si];
    McbSlot *slot = slots + slot_of[si];
    Lcg r{((uint64_t)(seed + nn) ^ 0x5DEECE66DULL) & LCG_MASK};
    auto draw = [&](int b) { return lcg_bound(r, b); };
    for (int st = 0; st < 2; st++) {
        const int zero = -N.first;
        int kzero = -1;
        for (int k = 0; k < N.nlev; k++) if (N.lvl[k] == zero) kzero = k;
        {
            double xo = lcg_double(r) * 256.0, yo = lcg_double(r) * 256.0, zo = lcg_double(r) * 256.0;
            fisher_yates(tile, lane, draw);
            if (kzero >= 0 && zero < N.octs) store_octave(slot, W->oct_base[nn] + st * N.nlev + kzero, tile, lane, xo, yo, zo);
        }
        for (int i = zero - 1; i >= 0; i--) {
            int kk = -1;
            for (int k = 0; k < N.nlev; k++) if (N.lvl[k] == i) kk = k;
            if (i < N.octs && kk >= 0) {
                double xo = lcg_double(r) * 256.0, yo = lcg_double(r) * 256.0, zo = lcg_double(r) * 256.0;
                fisher_yates(tile, lane, draw);
                store_octave(slot, W->oct_base[nn] + st * N.nlev + kk, tile, lane, xo, yo, zo);
            } else {
                r.s = (r.s * W->jm262 + W->ja262) & LCG_MASK;
            }
        }
    }
    if (nn == 0) slot->seed = seed;
}

// End: java.util.Random(seed), consumeCount(17292), SimplexNoise permutation (engine/mc_end.h:18-27)
__global__ __launch_bounds__(64) void k_seed_end(const McbWorld *__restrict__ W, const int64_t *__restrict__ seeds,
                                                const int32_t *__restrict__ slot_of, int n, McbSlot *__restrict__ slots) {
    __shared__ uint32_t tile[64 * 64];
    const int lane = threadIdx.x;
    const int gid = blockIdx.x * 64 + lane;
    if (gid >= n) return;
    const int64_t seed = seeds[gid];
    Lcg r{((uint64_t)seed ^ 0x5DEECE66DULL) & LCG_MASK};
    r.s = (r.s * W->jm_end + W->ja_end) & LCG_MASK;
    double xo = lcg_double(r) * 256.0, yo = lcg_double(r) * 256.0, zo = lcg_double(r) * 256.0;
    fisher_yates(tile, lane, [&](int b) { return lcg_bound(r, b); });
    McbSlot *slot = slots + slot_of[gid];
    store_octave(slot, 0, tile, lane, xo, yo, zo);
    slot->seed = seed;
}

// ================================================================================================
// grid kernel: tiles of quart-grid requests of one world
// ================================================================================================
// One workgroup = one tile of tile_w x (256/tile_w * cpl) columns; lane (lx, lz) evaluates the cpl
// columns (lx, lz*cpl + c) in order, each column's search warm-started from the previous one.
template <int MODE, int DIM, int MAXL, bool WRAP>
__global__ __launch_bounds__(256) void k_grid(const McbWorld *__restrict__ W, const McbNode *__restrict__ tree,
                                              const uint32_t *__restrict__ cells, const McbNode *__restrict__ crecs,
                                              const uint32_t *__restrict__ ucells, const McbNode *__restrict__ urecs,
                                              const McbSlot *__restrict__ slots, const McbRequest *__restrict__ reqs,
                                              const McbTile *__restrict__ tiles, int ntiles, int tile_w, int cpl,
                                              uint8_t *__restrict__ out) {
    __shared__ __attribute__((aligned(16))) uint8_t perm[(DIM == MCB_OVERWORLD ? MCB_OW_OCT : MCB_NE_OCT) * OCT_BYTES];
    __shared__ __attribute__((aligned(16))) typename GT<MODE>::T G[256];
    __shared__ uint32_t clut[DIM == MCB_OVERWORLD ? MCB_CELL_LUT : 1];
    __shared__ SplLds spl[1];
    SplLds *SP = spl;
    const int n_oct = DIM == MCB_OVERWORLD ? MCB_OW_OCT : MCB_NE_OCT;
    const int tree_n = W->tree_n;
    init_grad(G);                                         // all visible after the first staging barrier
    if (DIM == MCB_OVERWORLD) {
        stage_lut(clut, W);
        stage_spline(SP, W->spline);
    }
    const int tile_h = 256 / tile_w;
    int staged = -1;
    // contiguous tile range per workgroup, so consecutive tiles of one seed reuse the staged tables; the
    // XCD-aware rank puts the tiles of one seed on one L2, so all but the first stage its state from there
    const int nwg = (int)gridDim.x;
    const int vb = xcd_rank((int)blockIdx.x, nwg);
    const int per = (ntiles + nwg - 1) / nwg;
    const int t_end = min(ntiles, (vb + 1) * per);
    for (int ti = vb * per; ti < t_end; ti++) {
        const McbTile tl = tiles[ti];
        const McbRequest &R = reqs[tl.req];
        const int slot = R.slot;
        if (slot != staged) {
            __syncthreads();
            stage_perm2(perm, slots + slot, n_oct);
            __syncthreads();
            staged = slot;
        }
        const McbOct *oct = slots[slot].oct;
        const int ix = tl.tx * tile_w + (int)threadIdx.x % tile_w;
        const int iz0 = (tl.tz * tile_h + (int)threadIdx.x / tile_w) * cpl;
        if (ix >= R.nx || iz0 >= R.nz) continue;
        const int qx = R.x0 + ix;
        const int64_t plane = (int64_t)R.nx * R.nz;
        const int ny = R.ny;
        int cand[MAXL];
#pragma unroll
        for (int j = 0; j < MAXL; j++) cand[j] = -1;
        const int iz_end = min(iz0 + cpl, R.nz);
        for (int iz = iz0; iz < iz_end; iz++) {
            const int qz = R.z0 + iz;
            uint8_t *o = out + R.out_off + (int64_t)iz * R.nx + ix;
            if (DIM == MCB_OVERWORLD) {
                Column col;
                ow_column<MODE, const uint8_t *, WRAP>(W, perm, G, SP, oct, qx, qz, col);
                const int t5[5] = {col.t, col.h, col.c, col.e, col.w};
                int dep[MAXL], b[MAXL];
                bool in = W->cell_on && cell_in_range(t5[0]) && cell_in_range(t5[1]) && cell_in_range(t5[2]) &&
                          cell_in_range(t5[3]) && cell_in_range(t5[4]);
#pragma unroll
                for (int j = 0; j < MAXL; j++) {
                    dep[j] = j < ny ? ow_depth<MODE>(W, col, R.y[j]) : 0;
                    in = in && cell_in_range(dep[j]);
                }
                if (in) in = cell_layers<MAXL>(W, clut, cells, crecs, ucells, urecs, t5, dep, ny, b);
                if (!in) tree_search_layers<MAXL>(tree, tree_n, t5, dep, ny, cand, b);
#pragma unroll
                for (int j = 0; j < MAXL; j++)
                    if (j < ny) o[j * plane] = (uint8_t)b[j];
            } else {
                int t[6];
                ne_column<MODE, const uint8_t *, WRAP>(W, perm, G, oct, qx, qz, t);
                uint8_t b = (uint8_t)tree_search(tree, tree_n, t, cand[0]);
                for (int j = 0; j < ny; j++) o[j * plane] = b;
            }
        }
    }
}

// End tiles.  The biome of a quart depends only on its chunk, through the island field of the 25x25
// cells around the chunk (engine/mc_end.h:66-119).  A tile first finds the islands of the union of its
// chunks' cell windows (one simplex evaluation per cell, spread over the workgroup, islands appended to
// an LDS list), then every lane takes the maximum over the listed islands inside its chunk's window.  The
// maximum does not depend on the order of the cells, so the result is the reference's.
template <int MODE>
__global__ __launch_bounds__(256) void k_grid_end(const McbWorld *__restrict__ W, const McbSlot *__restrict__ slots,
                                                  const McbRequest *__restrict__ reqs, const McbTile *__restrict__ tiles,
                                                  int ntiles, int tile_w, int cpl, uint8_t *__restrict__ out) {
    __shared__ __attribute__((aligned(16))) uint8_t perm[256];
    __shared__ int isl_x[END_CAP], isl_z[END_CAP];
    __shared__ float isl_s[END_CAP];
    __shared__ int isl_n;
    const int tile_h = 256 / tile_w;
    int staged = -1;
    const int vb = xcd_rank((int)blockIdx.x, (int)gridDim.x);      // tiles of one seed on one L2
    const int per = (ntiles + (int)gridDim.x - 1) / (int)gridDim.x;
    const int t_end = min(ntiles, (vb + 1) * per);
    for (int ti = vb * per; ti < t_end; ti++) {
        const McbTile tl = tiles[ti];
        const McbRequest &R = reqs[tl.req];
        __syncthreads();                                  // previous tile's list fully consumed
        if (R.slot != staged) {
            stage_perm(perm, slots + R.slot, 1);
            staged = R.slot;
        }
        if (threadIdx.x == 0) isl_n = 0;
        // quart, chunk and island-cell extent of this tile
        const int ix_lo = tl.tx * tile_w, iz_lo = tl.tz * tile_h * cpl;
        const int ix_hi = min(ix_lo + tile_w, R.nx) - 1, iz_hi = min(iz_lo + tile_h * cpl, R.nz) - 1;
        const int cx_lo = (R.x0 + ix_lo) >> 2, cx_hi = (R.x0 + ix_hi) >> 2;
        const int cz_lo = (R.z0 + iz_lo) >> 2, cz_hi = (R.z0 + iz_hi) >> 2;
        // chunk c is sampled at section 2c+1, whose island cells are (2c+1)/2 + [-12, 12]
        const int wx_lo = (2 * cx_lo + 1) / 2 - 12, wx_hi = (2 * cx_hi + 1) / 2 + 12;
        const int wz_lo = (2 * cz_lo + 1) / 2 - 12, wz_hi = (2 * cz_hi + 1) / 2 + 12;
        const int wxn = wx_hi - wx_lo + 1, ncell = wxn * (wz_hi - wz_lo + 1);
        __syncthreads();
        for (int c = (int)threadIdx.x; c < ncell; c += 256) {
            const int64_t tx = wx_lo + c % wxn, tz = wz_lo + c / wxn;
            if (tx * tx + tz * tz <= 4096) continue;
            const double nv = simplex2(perm, (double)tx, (double)tz);
            const bool island = MODE ? ((float)nv < -0.9f) : (nv < (double)-0.9f);
            if (!island) continue;
            const int k = atomicAdd(&isl_n, 1);
            if (k < END_CAP) {
                isl_x[k] = (int)tx;
                isl_z[k] = (int)tz;
                isl_s[k] = fmodf(fabsf((float)tx) * 3439.0f + fabsf((float)tz) * 147.0f, 13.0f) + 9.0f;
            }
        }
        __syncthreads();
        const int n = isl_n;
        const int ix = ix_lo + (int)threadIdx.x % tile_w;
        const int iz0 = (tl.tz * tile_h + (int)threadIdx.x / tile_w) * cpl;
        if (ix >= R.nx || iz0 >= R.nz) continue;
        const int64_t plane = (int64_t)R.nx * R.nz;
        const int iz_end = min(iz0 + cpl, R.nz);
        for (int iz = iz0; iz < iz_end; iz++) {
            const int qx = R.x0 + ix, qz = R.z0 + iz;
            const int cx = qx >> 2, cz = qz >> 2;
            int b;
            if ((int64_t)cx * cx + (int64_t)cz * cz <= 4096) b = W->end_ids[0];
            else if (n > END_CAP) b = end_biome<MODE>(W, perm, qx, qz);      // list overflow: direct scan
            else b = end_biome_from_height<MODE>(W, cx, cz, end_height_list<MODE>(isl_x, isl_z, isl_s, n, 2 * cx + 1, 2 * cz + 1));
            for (int j = 0; j < R.ny; j++) out[R.out_off + j * plane + (int64_t)iz * R.nx + ix] = (uint8_t)b;
        }
    }
}

// ================================================================================================
// End island field over a block grid of every seed (mc_end.h:66-106): kind 0 = the height value at
// section (bx/8, bz/8), kind 1 = the End's erosion value at block (bx, bz).  A 16x16 tile lists the
// islands of the union of its sections' cell windows once (as k_grid_end); a tile whose window is
// larger than END_WINDOW cells, or whose list overflows, evaluates its samples directly.
// ================================================================================================
#define END_WINDOW 16384
template <int MODE>
__global__ __launch_bounds__(256) void k_end_field(const McbSlot *__restrict__ slots, int tiles_x, int bx0, int bz0, int nx,
                                                   int nz, int step, int kind, double *__restrict__ out) {
    __shared__ __attribute__((aligned(16))) uint8_t perm[256];
    __shared__ int isl_x[END_CAP], isl_z[END_CAP];
    __shared__ float isl_s[END_CAP];
    __shared__ int isl_n;
    const int s = (int)blockIdx.y;
    const int tx = (int)blockIdx.x % tiles_x, tz = (int)blockIdx.x / tiles_x;
    stage_perm(perm, slots + s, 1);
    if (threadIdx.x == 0) isl_n = 0;
    const int ix_lo = tx * 16, iz_lo = tz * 16;
    const int ix_hi = min(ix_lo + 16, nx) - 1, iz_hi = min(iz_lo + 16, nz) - 1;
    // chunk of a block = (b / 8) / 2 with C division: monotonic in b, so the tile's chunks form a range
    const int cx_lo = (bx0 + ix_lo * step) / 8 / 2, cx_hi = (bx0 + ix_hi * step) / 8 / 2;
    const int cz_lo = (bz0 + iz_lo * step) / 8 / 2, cz_hi = (bz0 + iz_hi * step) / 8 / 2;
    const int wx_lo = cx_lo - 12, wz_lo = cz_lo - 12;
    const int64_t wxn = (int64_t)cx_hi - cx_lo + 25, wzn = (int64_t)cz_hi - cz_lo + 25;
    const bool listed = wxn * wzn <= END_WINDOW;
    __syncthreads();
    if (listed) {
        const int ncell = (int)(wxn * wzn);
        for (int c = (int)threadIdx.x; c < ncell; c += 256) {
            const int64_t cx = wx_lo + c % (int)wxn, cz = wz_lo + c / (int)wxn;
            if (cx * cx + cz * cz <= 4096) continue;
            const double nv = simplex2(perm, (double)cx, (double)cz);
            const bool island = MODE ? ((float)nv < -0.9f) : (nv < (double)-0.9f);
            if (!island) continue;
            const int k = atomicAdd(&isl_n, 1);
            if (k < END_CAP) {
                isl_x[k] = (int)cx;
                isl_z[k] = (int)cz;
                isl_s[k] = fmodf(fabsf((float)cx) * 3439.0f + fabsf((float)cz) * 147.0f, 13.0f) + 9.0f;
            }
        }
    }
    __syncthreads();
    const int n = isl_n;
    const int ix = ix_lo + (int)threadIdx.x % 16, iz = iz_lo + (int)threadIdx.x / 16;
    if (ix >= nx || iz >= nz) return;
    const int bx = bx0 + ix * step, bz = bz0 + iz * step;
    const int sx = bx / 8, sz = bz / 8;
    const float h = (listed && n <= END_CAP) ? end_height_list<MODE>(isl_x, isl_z, isl_s, n, sx, sz)
                                             : end_height<MODE>(perm, sx, sz);
    out[(size_t)s * nx * nz + (size_t)iz * nx + ix] = kind == 0 ? (double)h : end_erosion_from_height<MODE>(bx, bz, h);
}

// ================================================================================================
// points kernels: seeds x quart points (mcg_ref_points contract, engine/cuda/mcg_host.c:36-83)
// ================================================================================================
// Biome (and targets) of one quart point of one seed.  PS = where the permutation tables are (the
// doubled LDS copy, or the seed slot itself); End: P = the seed's 256-byte island table.
template <int MODE, int DIM, class PS>
__device__ __forceinline__ int point_biome(const McbWorld *__restrict__ W, const McbNode *__restrict__ tree,
                                           const uint32_t *__restrict__ cells, const McbNode *__restrict__ crecs, PS perm,
                                           const uint8_t *P, const typename GT<MODE>::T *G, const SplLds *SP,
                                           const uint32_t *clut, const McbOct *oct, int qx, int qy, int qz, int *t) {
    int b;
    if constexpr (DIM == MCB_OVERWORLD) {
        Column c;
        ow_column<MODE>(W, perm, G, SP, oct, qx, qz, c);
        t[0] = c.t; t[1] = c.h; t[2] = c.c; t[3] = c.e; t[5] = c.w;
        t[4] = ow_depth<MODE>(W, c, qy);
        const int t5[5] = {c.t, c.h, c.c, c.e, c.w};
        uint32_t e = 0;
        if (W->cell_on && cell_in_range(t5[0]) && cell_in_range(t5[1]) && cell_in_range(t5[2]) && cell_in_range(t5[3]) &&
            cell_in_range(t5[4]) && cell_in_range(t[4]))
            e = cells[cell_base(W, clut, t5) + cell_depth(clut, t[4])];
        if (e & 255u) {
            b = cell_best(crecs, e, t5, t[4]);
        } else {
            int cand = -1;
            b = tree_search(tree, W->tree_n, t, cand);
        }
    } else if constexpr (DIM == MCB_NETHER) {
        ne_column<MODE>(W, perm, G, oct, qx, qz, t);
        int cand = -1;
        b = tree_search(tree, W->tree_n, t, cand);
    } else {
        b = end_biome<MODE>(W, P, qx, qz);
    }
    return b;
}

// Many points per seed: a workgroup per (seed, 256 points), the seed's tables staged in LDS.
template <int MODE, int DIM>
__global__ __launch_bounds__(256) void k_points(const McbWorld *__restrict__ W, const McbNode *__restrict__ tree,
                                                const uint32_t *__restrict__ cells, const McbNode *__restrict__ crecs,
                                                const McbSlot *__restrict__ slots, const int32_t *__restrict__ slot_of,
                                                const int32_t *__restrict__ pts, int np, uint8_t *__restrict__ out,
                                                int32_t *__restrict__ tg) {
    __shared__ __attribute__((aligned(16))) uint8_t perm[(DIM == MCB_OVERWORLD ? MCB_OW_OCT : MCB_NE_OCT) * OCT_BYTES];
    __shared__ __attribute__((aligned(16))) typename GT<MODE>::T G[256];
    const int n_oct = DIM == MCB_OVERWORLD ? MCB_OW_OCT : (DIM == MCB_NETHER ? MCB_NE_OCT : 1);
    const int s = blockIdx.y;
    const int slot = slot_of[s];
    __shared__ uint32_t clut[DIM == MCB_OVERWORLD ? MCB_CELL_LUT : 1];
    __shared__ SplLds spl[1];
    if (DIM == MCB_END) stage_perm(perm, slots + slot, 1);
    else stage_perm2(perm, slots + slot, n_oct);
    init_grad(G);
    if (DIM == MCB_OVERWORLD) {
        stage_lut(clut, W);
        stage_spline(spl, W->spline);
    }
    __syncthreads();
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= np) return;
    const size_t o = (size_t)s * np + p;
    int t[6] = {0, 0, 0, 0, 0, 0};
    const int b = point_biome<MODE, DIM>(W, tree, cells, crecs, (const uint8_t *)perm, perm, G, spl, clut, slots[slot].oct,
                                         pts[3 * p], pts[3 * p + 1], pts[3 * p + 2], t);
    out[o] = (uint8_t)b;
    if (tg)
        for (int k = 0; k < 6; k++) tg[o * 6 + k] = t[k];
}

// Few points per seed: a lane per (seed, point), grid-stride; the permutations and octave records are
// read from the seed slot in global memory (masked), and only the seed-independent tables (gradients,
// cell locator, spline) are staged, once per workgroup.  Same arithmetic as k_points.
template <int MODE, int DIM>
__global__ __launch_bounds__(256) void k_points_g(const McbWorld *__restrict__ W, const McbNode *__restrict__ tree,
                                                  const uint32_t *__restrict__ cells, const McbNode *__restrict__ crecs,
                                                  const McbSlot *__restrict__ slots, const int32_t *__restrict__ slot_of, int ns,
                                                  const int32_t *__restrict__ pts, int np, uint8_t *__restrict__ out,
                                                  int32_t *__restrict__ tg) {
    __shared__ __attribute__((aligned(16))) typename GT<MODE>::T G[256];
    __shared__ uint32_t clut[DIM == MCB_OVERWORLD ? MCB_CELL_LUT : 1];
    __shared__ SplLds spl[1];
    init_grad(G);
    if (DIM == MCB_OVERWORLD) {
        stage_lut(clut, W);
        stage_spline(spl, W->spline);
    }
    __syncthreads();
    const int64_t n = (int64_t)ns * np;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const int s = (int)(i / np), p = (int)(i - (int64_t)s * np);
        const McbSlot *sl = slots + slot_of[s];
        int t[6] = {0, 0, 0, 0, 0, 0};
        const int b = point_biome<MODE, DIM>(W, tree, cells, crecs, SlotPerm{sl->perm}, sl->perm[0], G, spl, clut, sl->oct,
                                             pts[3 * p], pts[3 * p + 1], pts[3 * p + 2], t);
        out[i] = (uint8_t)b;
        if (tg)
            for (int k = 0; k < 6; k++) tg[i * 6 + k] = t[k];
    }
}

// ================================================================================================
// biomes of quantized climate targets {t,h,c,e,d,w} (the search step alone, engine/mc_biomes.h:31-62)
// ================================================================================================
template <int DIM>
__global__ __launch_bounds__(256) void k_targets(const McbWorld *__restrict__ W, const McbNode *__restrict__ tree,
                                                 const uint32_t *__restrict__ cells, const McbNode *__restrict__ crecs,
                                                 const int32_t *__restrict__ tg, int n, uint8_t *__restrict__ out) {
    const int i = (int)blockIdx.x * 256 + (int)threadIdx.x;
    if (i >= n) return;
    int t[6];
#pragma unroll
    for (int k = 0; k < 6; k++) t[k] = tg[6 * i + k];
    int b;
    uint32_t e = 0;
    const int t5[5] = {t[0], t[1], t[2], t[3], t[5]};
    if (DIM == MCB_OVERWORLD && W->cell_on && cell_in_range(t[0]) && cell_in_range(t[1]) && cell_in_range(t[2]) &&
        cell_in_range(t[3]) && cell_in_range(t[4]) && cell_in_range(t[5]))
        e = cells[cell_base(W, W->cell_lut[0], t5) + cell_depth(W->cell_lut[0], t[4])];
    if (e & 255u) {
        b = cell_best(crecs, e, t5, t[4]);
    } else {
        int cand = -1;
        b = tree_search(tree, W->tree_n, t, cand);
    }
    out[i] = (uint8_t)b;
}

// ================================================================================================
// BiomeManager zoom (engine/cuda/mc_gpu.cuh:96-118, libmcgen/src/region.c:62-85): block -> one of the
// 8 surrounding quart cells by the smallest fiddled distance; reads the quart layers of its request.
// ================================================================================================
__global__ __launch_bounds__(256) void k_zoom(const McbRequest *__restrict__ zreqs, const McbRequest *__restrict__ qreqs,
                                              const McbTile *__restrict__ tiles, int ntiles, const uint8_t *__restrict__ quart,
                                              uint8_t *__restrict__ out) {
    // offsets of the tile's lattice points, per layer: a 16 x 16 tile spans at most 6 x 6 corners in x, z
    __shared__ ZFid F[8 * 2 * 6 * 6];
    // neighbouring tiles of a request (rows of one output line) on one L2
    for (int ti = xcd_rank((int)blockIdx.x, (int)gridDim.x); ti < ntiles; ti += gridDim.x) {
        const McbTile tl = tiles[ti];
        const McbRequest &Z = zreqs[tl.req];
        const McbRequest &Q = qreqs[tl.req];
        const int hx0 = tl.tx * 16, hz0 = tl.tz * 16;
        const int hx1 = min(hx0 + 16, Z.nx) - 1, hz1 = min(hz0 + 16, Z.nz) - 1;
        const int pxl = (Z.x0 + hx0 - 2) >> 2, npx = ((Z.x0 + hx1 - 2) >> 2) + 2 - pxl;
        const int pzl = (Z.z0 + hz0 - 2) >> 2, npz = ((Z.z0 + hz1 - 2) >> 2) + 2 - pzl;
        const int npl = npx * npz;
        const int64_t hs = Z.hashed_seed;
        __syncthreads();                                  // the previous tile's readers are done
        for (int i = threadIdx.x; i < Z.ny * 2 * npl; i += 256) {
            const int u = i % npx, v = (i / npx) % npz, w = i / npl;      // w = 2 * layer + corner y
            F[i] = zoom_fiddles(hs, pxl + u, ((Z.y[w >> 1] - 2) >> 2) + (w & 1), pzl + v);
        }
        __syncthreads();
        const int ix = hx0 + (int)threadIdx.x % 16, iz = hz0 + (int)threadIdx.x / 16;
        if (ix > hx1 || iz > hz1) continue;
        const int x = Z.x0 + ix, z = Z.z0 + iz;
        const int64_t qplane = (int64_t)Q.nx * Q.nz;
        for (int j = 0; j < Z.ny; j++) {
            const int py = (Z.y[j] - 2) >> 2;
            int q[3];
            zoom_pick(x, Z.y[j], z, [&](int, int cx, int cy, int cz) {
                return F[((2 * j + (cy - py)) * npz + (cz - pzl)) * npx + (cx - pxl)];
            }, q);
            const int qxi = q[0] - Q.x0, qzi = q[2] - Q.z0;
            const int layer = 2 * j + (q[1] - py);
            out[Z.out_off + j * (int64_t)Z.nx * Z.nz + (int64_t)iz * Z.nx + ix] =
                quart[Q.out_off + layer * qplane + (int64_t)qzi * Q.nx + qxi];
        }
    }
}

template <int MODE, int DIM, int MAXL, bool WRAP>
hipError_t grid_launch3(const DevWorld &w, const McbSlot *slots, const McbRequest *reqs, const McbTile *tiles, int ntiles,
                        uint8_t *out, int tile_w, int cpl, int blocks, hipStream_t st) {
    if constexpr (DIM == MCB_END)
        hipLaunchKernelGGL((k_grid_end<MODE>), dim3(blocks), dim3(256), 0, st, w.d_world, slots, reqs, tiles, ntiles, tile_w,
                           cpl, out);
    else
        hipLaunchKernelGGL((k_grid<MODE, DIM, MAXL, WRAP>), dim3(blocks), dim3(256), 0, st, w.d_world, w.d_tree, w.d_cells,
                           w.d_cell_recs, w.d_cell_u, w.d_cell_urecs, slots, reqs, tiles, ntiles, tile_w, cpl, out);
    return hipGetLastError();
}
// The grid kernels have a variant without PerlinNoise.wrap's range test, for launches whose octave inputs
// provably stay inside +-2^24 (wrap_free); the test and its branch cost about 9% of the 26.3 kernel
// (26.3 grid 2.84 -> 2.59 ms).  26.1/26.2 take it up to two layers per column; at four the register
// allocator gives it 132 VGPRs (3 waves per SIMD) and it measured slower (3.52 -> 3.69 ms).
template <int MODE, int DIM, int MAXL>
hipError_t grid_launch2(const DevWorld &w, const McbSlot *slots, const McbRequest *reqs, const McbTile *tiles, int ntiles,
                        uint8_t *out, int tile_w, int cpl, int blocks, bool nowrap, hipStream_t st) {
    if constexpr (DIM != MCB_END && (MODE == 1 || MAXL <= 2))
        if (nowrap) return grid_launch3<MODE, DIM, MAXL, false>(w, slots, reqs, tiles, ntiles, out, tile_w, cpl, blocks, st);
    return grid_launch3<MODE, DIM, MAXL, true>(w, slots, reqs, tiles, ntiles, out, tile_w, cpl, blocks, st);
}
template <int MODE, int DIM>
hipError_t grid_launch(const DevWorld &w, const McbSlot *slots, const McbRequest *reqs, const McbTile *tiles, int ntiles,
                       uint8_t *out, int tile_w, int cpl, int max_ny, int blocks, bool nowrap, hipStream_t st) {
    if constexpr (DIM != MCB_OVERWORLD) {
        return grid_launch2<MODE, DIM, 1>(w, slots, reqs, tiles, ntiles, out, tile_w, cpl, blocks, nowrap, st);
    } else {
        if (max_ny <= 1) return grid_launch2<MODE, DIM, 1>(w, slots, reqs, tiles, ntiles, out, tile_w, cpl, blocks, nowrap, st);
        if (max_ny <= 2) return grid_launch2<MODE, DIM, 2>(w, slots, reqs, tiles, ntiles, out, tile_w, cpl, blocks, nowrap, st);
        if (max_ny <= 4) return grid_launch2<MODE, DIM, 4>(w, slots, reqs, tiles, ntiles, out, tile_w, cpl, blocks, nowrap, st);
        return grid_launch2<MODE, DIM, 8>(w, slots, reqs, tiles, ntiles, out, tile_w, cpl, blocks, nowrap, st);
    }
}
// Points per seed from which staging the seed in LDS (k_points) beats reading its slot (k_points_g),
// measured on MI355X over 16..1024 points per seed: Overworld between 64 and 128 (26.3 and 26.1 alike),
// Nether between 32 and 64, End between 128 and 256.
constexpr int points_stage_min(int dim) { return dim == MCB_OVERWORLD ? 128 : (dim == MCB_NETHER ? 32 : 256); }
template <int MODE, int DIM>
hipError_t points_launch(const DevWorld &w, const McbSlot *slots, const int32_t *slot_of, int ns, const int32_t *pts, int np,
                         uint8_t *out, int32_t *tg, hipStream_t st) {
    if (np < points_stage_min(DIM)) {
        const int64_t n = (int64_t)ns * np;
        const int blocks = (int)std::min<int64_t>((n + 255) / 256, 4096);
        hipLaunchKernelGGL((k_points_g<MODE, DIM>), dim3(blocks), dim3(256), 0, st, w.d_world, w.d_tree, w.d_cells,
                           w.d_cell_recs, slots, slot_of, ns, pts, np, out, tg);
    } else {
        hipLaunchKernelGGL((k_points<MODE, DIM>), dim3((np + 255) / 256, ns), dim3(256), 0, st, w.d_world, w.d_tree, w.d_cells,
                           w.d_cell_recs, slots, slot_of, pts, np, out, tg);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_seed_init(const DevWorld &w, const int64_t *seeds, const int32_t *slot_of, int n, McbSlot *slots,
                            hipStream_t st) {
    if (n <= 0) return hipSuccess;
    int lanes = w.dim == MCB_OVERWORLD ? n * w.host.n_oct : (w.dim == MCB_NETHER ? n * 2 : n);
    dim3 grid((lanes + 63) / 64);
    if (w.dim == MCB_OVERWORLD) hipLaunchKernelGGL(k_seed_ow, grid, dim3(64), 0, st, w.d_world, seeds, slot_of, n, slots);
    else if (w.dim == MCB_NETHER) hipLaunchKernelGGL(k_seed_ne, grid, dim3(64), 0, st, w.d_world, seeds, slot_of, n, slots);
    else hipLaunchKernelGGL(k_seed_end, grid, dim3(64), 0, st, w.d_world, seeds, slot_of, n, slots);
    return hipGetLastError();
}

hipError_t launch_grid(const DevWorld &w, const McbSlot *slots, const McbRequest *reqs, const McbTile *tiles, int ntiles,
                       uint8_t *out, int tile_w, int cpl, int max_ny, int blocks, bool nowrap, hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    if (tile_w <= 0 || tile_w > 256 || 256 % tile_w || cpl < 1 || cpl > 64 || max_ny < 1 || max_ny > 8)
        return hipErrorInvalidValue;
    if (blocks <= 0) blocks = ntiles;
    const bool f = w.host.fmt_new != 0;
    switch (w.dim) {
        case MCB_OVERWORLD:
            return f ? grid_launch<1, MCB_OVERWORLD>(w, slots, reqs, tiles, ntiles, out, tile_w, cpl, max_ny, blocks, nowrap, st)
                     : grid_launch<0, MCB_OVERWORLD>(w, slots, reqs, tiles, ntiles, out, tile_w, cpl, max_ny, blocks, nowrap, st);
        case MCB_NETHER:
            return f ? grid_launch<1, MCB_NETHER>(w, slots, reqs, tiles, ntiles, out, tile_w, cpl, max_ny, blocks, nowrap, st)
                     : grid_launch<0, MCB_NETHER>(w, slots, reqs, tiles, ntiles, out, tile_w, cpl, max_ny, blocks, nowrap, st);
        default:
            return f ? grid_launch<1, MCB_END>(w, slots, reqs, tiles, ntiles, out, tile_w, cpl, max_ny, blocks, nowrap, st)
                     : grid_launch<0, MCB_END>(w, slots, reqs, tiles, ntiles, out, tile_w, cpl, max_ny, blocks, nowrap, st);
    }
}

hipError_t launch_points(const DevWorld &w, const McbSlot *slots, const int32_t *slot_of, int ns, const int32_t *pts, int np,
                         uint8_t *out, int32_t *tg, hipStream_t st) {
    if (ns <= 0 || np <= 0) return hipSuccess;
    const bool f = w.host.fmt_new != 0;
    switch (w.dim) {
        case MCB_OVERWORLD: return f ? points_launch<1, MCB_OVERWORLD>(w, slots, slot_of, ns, pts, np, out, tg, st)
                                     : points_launch<0, MCB_OVERWORLD>(w, slots, slot_of, ns, pts, np, out, tg, st);
        case MCB_NETHER: return f ? points_launch<1, MCB_NETHER>(w, slots, slot_of, ns, pts, np, out, tg, st)
                                  : points_launch<0, MCB_NETHER>(w, slots, slot_of, ns, pts, np, out, tg, st);
        default: return f ? points_launch<1, MCB_END>(w, slots, slot_of, ns, pts, np, out, tg, st)
                          : points_launch<0, MCB_END>(w, slots, slot_of, ns, pts, np, out, tg, st);
    }
}

hipError_t launch_targets(const DevWorld &w, const int32_t *tg, int n, uint8_t *out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (w.dim == MCB_OVERWORLD)
        hipLaunchKernelGGL(k_targets<MCB_OVERWORLD>, dim3((n + 255) / 256), dim3(256), 0, st, w.d_world, w.d_tree, w.d_cells,
                           w.d_cell_recs, tg, n, out);
    else if (w.dim == MCB_NETHER)
        hipLaunchKernelGGL(k_targets<MCB_NETHER>, dim3((n + 255) / 256), dim3(256), 0, st, w.d_world, w.d_tree, w.d_cells,
                           w.d_cell_recs, tg, n, out);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// Tie masks by brute force over every leaf (engine/cuda/mcg_host.c:85-96): one workgroup per target,
// lanes stride over the pre-order node array; a leaf's distance is the reference's 64-bit integer sum of
// per-dimension squared interval distances (wrapping like the reference's i64 on overflow), compared
// signed.  Each lane keeps (least distance, biome mask); the workgroup folds them in LDS: the smaller
// distance wins, equal distances OR their masks.
__global__ __launch_bounds__(256) void k_tie_masks(const McbNode *__restrict__ N, int n, const int32_t *__restrict__ tg,
                                                   uint64_t *__restrict__ out) {
    __shared__ int64_t sd[256];
    __shared__ uint64_t s0[256], s1[256];
    const int lane = (int)threadIdx.x;
    const int32_t *t = tg + 6 * (int64_t)blockIdx.x;
    int64_t t6[6];
#pragma unroll
    for (int k = 0; k < 6; k++) t6[k] = t[k];
    int64_t best = INT64_MAX;
    uint64_t m0 = 0, m1 = 0;
    for (int j = lane; j < n; j += 256) {
        const McbNode nd = N[j];
        if (!nd.leaf) continue;
        uint64_t d = nd.d6;
#pragma unroll
        for (int k = 0; k < 6; k++) {
            const int64_t lo = nd.lo[k], hi = nd.hi[k];
            const int64_t x = t6[k] < lo ? lo - t6[k] : (t6[k] > hi ? t6[k] - hi : 0);
            d += (uint64_t)x * (uint64_t)x;
        }
        const int64_t ds = (int64_t)d;
        if (ds < best) { best = ds; m0 = m1 = 0; }
        if (ds == best) {
            const int b = nd.leaf & 0xFF;
            if (b < 64) m0 |= 1ULL << b; else m1 |= 1ULL << (b - 64);
        }
    }
    sd[lane] = best;
    s0[lane] = m0;
    s1[lane] = m1;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (lane < w) {
            const int64_t a = sd[lane], b = sd[lane + w];
            if (b < a) { sd[lane] = b; s0[lane] = s0[lane + w]; s1[lane] = s1[lane + w]; }
            else if (b == a) { s0[lane] |= s0[lane + w]; s1[lane] |= s1[lane + w]; }
        }
        __syncthreads();
    }
    if (lane == 0) {
        out[2 * (int64_t)blockIdx.x] = s0[0];
        out[2 * (int64_t)blockIdx.x + 1] = s1[0];
    }
}

hipError_t launch_tie_masks(const DevWorld &w, const int32_t *tg, int n, uint64_t *out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tie_masks, dim3(n), dim3(256), 0, st, w.d_tree, w.host.tree_n, tg, out);
    return hipGetLastError();
}

hipError_t launch_end_field(const DevWorld &w, const McbSlot *slots, int ns, int bx0, int bz0, int nx, int nz, int step,
                            int kind, double *out, hipStream_t st) {
    if (ns <= 0 || nx <= 0 || nz <= 0) return hipSuccess;
    if (w.dim != MCB_END || ns > 65535) return hipErrorInvalidValue;
    const int tiles_x = (nx + 15) / 16, tiles = tiles_x * ((nz + 15) / 16);
    if (w.host.fmt_new)
        hipLaunchKernelGGL(k_end_field<1>, dim3(tiles, ns), dim3(256), 0, st, slots, tiles_x, bx0, bz0, nx, nz, step, kind, out);
    else
        hipLaunchKernelGGL(k_end_field<0>, dim3(tiles, ns), dim3(256), 0, st, slots, tiles_x, bx0, bz0, nx, nz, step, kind, out);
    return hipGetLastError();
}

hipError_t launch_zoom(const McbRequest *zreqs, const McbRequest *qreqs, const McbTile *tiles, int ntiles, const uint8_t *quart,
                       uint8_t *out, hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zoom, dim3(ntiles), dim3(256), 0, st, zreqs, qreqs, tiles, ntiles, quart, out);
    return hipGetLastError();
}

}  // namespace mcb
