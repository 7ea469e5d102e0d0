// mcg_plugin.hip - kernels of the libmcgen GPU plugin (include/mcgpu.h).
//
// One lane per quart cell (or per point): BiomeManager zoom -> climate program -> R-tree search with the
// chunk-fill-order tie chain.  Semantics follow the reference plugin and the libmcgen CPU code it mirrors
// (libmcgen/gpu/mcgpu_dev.cuh:1-544 and mcgpu_biome.cu:13-122; CPU: libmcgen/src/noise.c:339-347,
// df_new.c:655-836, df_old.c, biome.c:26-120), with the same order of every floating-point operation; the
// build uses -ffp-contract=off and no fast-math, so each multiply and add rounds separately.  The fused
// multiply-adds written below have an operand in {-1, 0, 1}: the product is exact, so they round like the
// separate product and sum (same value, same signed zero).
//
// MI355X layout: the compiled program (PIns) is read with scalar loads, every lane running the same
// instruction; live values sit in LDS slots, column per lane (slot s of lane l at s*256 + l, no bank
// conflicts); the octave permutation tables are staged in LDS once per workgroup when they fit
// (kPermDoubled chooses single or doubled tables), and a gradient table per permutation value (the
// engine's: f32/f16 entries for float noise, 16-byte double entries for the legacy path) replaces the
// bit-field extraction of every corner's gradient.
#include "mcg_plugin_dev.h"

namespace mcb {
namespace plug {

// ---------------------------------------------------------------------------------------------------
// R-tree (mcgpu_dev.cuh:440-500): distances in wrapping 64-bit arithmetic as the reference's u64 sums
// ---------------------------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t pdist(const PNode &n, const int64_t *tg) {
    uint64_t d = n.d6;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int64_t above = tg[k] - (int64_t)n.hi[k], below = (int64_t)n.lo[k] - tg[k];
        const int64_t x = above > 0 ? above : (below > 0 ? below : 0);
        d += (uint64_t)x * (uint64_t)x;
    }
    return d;
}
// first leaf of least distance in the reference's child order (its frame walk keeps one running best
// and enters a child when best > distance); -1 when none is accepted
__device__ int rt_find(const PNode *__restrict__ T, int n, const int64_t *tg) {
    uint64_t best = 0x7fffffffffffffffULL;
    int bl = -1, cur = 1;
    while (cur < n) {
        const PNode nd = T[cur];
        const uint64_t cd = pdist(nd, tg);
        if (best > cd) {
            if (nd.leaf) { best = cd; bl = cur; cur = nd.escape; }
            else cur++;
        } else {
            cur = nd.escape;
        }
    }
    return bl;
}
// at least two leaves at distance d (children entered when their distance is <= d)
__device__ bool rt_tie(const PNode *__restrict__ T, int n, const int64_t *tg, uint64_t d) {
    if (n < 2 || T[0].leaf) return false;
    int cnt = 0, cur = 1;
    while (cur < n) {
        const PNode nd = T[cur];
        const uint64_t cd = pdist(nd, tg);
        if (cd > d) { cur = nd.escape; continue; }
        if (nd.leaf) {
            if (cd == d && ++cnt >= 2) return true;
            cur = nd.escape;
        } else {
            cur++;
        }
    }
    return false;
}

// least-distance search through the cell table: the cell's list holds (in pre-order) every leaf that can
// be nearest or tie at the least distance there, so the first minimum of the list is the walk's result
// and the number of leaves at that distance is the tie count.  Within the table's +-2^17 range the
// distances stay far below 2^63, so no sum wraps.  Returns false when the target is outside the table
// or its cell is not listed (the caller walks the tree).
struct Found {
    int leaf, biome;               // the first leaf at the least distance and its biome
    uint64_t d;                    // the least distance (unset when single)
    bool tied;                     // two or more leaves at the least distance
    bool single;                   // the cell lists one leaf: nothing else can be at the least distance
};
// Targets within 16 bits: the cell offset of each dimension from the locator (one load per dimension
// instead of a binary search).  A cell with one listed leaf answers from its entry; otherwise the
// listed leaves' compact records are scanned with distances in doubles: every term is the square of an
// integer below 2^17 and there are seven, so each sum is an exact integer (the u64 sum's value).
__device__ __forceinline__ double rdist_d(const PRec &r, const int *t) {
    double d = (double)r.d6;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int x = max(max(t[k] - (int)r.hi[k], (int)r.lo[k] - t[k]), 0);
        d = fma((double)x, (double)x, d);
    }
    return d;
}
__device__ __forceinline__ bool cell_search_lut(const PWorld &W, const int *t, Found &f) {
    uint32_t idx = 0;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const uint32_t u = (uint32_t)(t[k] + 32768);
        const uint32_t e = W.clut[k * 256 + (int)(u >> 8)];
        idx += (e & 0xFFFFFu) + ((u & 255u) >= (e >> 20) ? (uint32_t)W.cstride[k] : 0u);
    }
    const uint32_t e = W.cent1[idx];
    const int cnt = (int)(e & 0xFF);
    if (!cnt) return false;
    f.tied = false;
    if (cnt == 1) {
        f.single = true;
        f.leaf = (int)((e >> 8) & 0xFFFFu);
        f.biome = (int)(e >> 24);
        return true;
    }
    f.single = false;
    const PRec *R = W.crec + (e >> 8);
    double best = __builtin_inf();
    int bl = 0, nt = 0;
    for (int q = 0; q < cnt; q++) {
        const PRec r = R[q];
        const double dq = rdist_d(r, t);
        if (dq < best) { best = dq; bl = r.leaf; nt = 1; }
        else if (dq == best) nt++;
    }
    f.leaf = bl & 0xFFFF;
    f.biome = (bl >> 16) & 0xFF;
    f.d = (uint64_t)best;
    f.tied = nt >= 2;
    return true;
}
__device__ __forceinline__ bool cell_search(const PWorld &W, const int64_t *tg, Found &f) {
    if (!W.centry) return false;
    if (W.clut) {
        bool in16 = true;
#pragma unroll
        for (int k = 0; k < 6; k++) in16 = in16 && tg[k] >= -32768 && tg[k] <= 32767;
        if (in16) {
            int t[6];
#pragma unroll
            for (int k = 0; k < 6; k++) t[k] = (int)tg[k];
            return cell_search_lut(W, t, f);
        }
    }
    int c[6];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int64_t t = tg[k];
        if (t < -(int64_t)kCellTMax || t > (int64_t)kCellTMax) return false;
        const int32_t *bp = W.cbp + W.coff[k];
        int lo = 0, hi = W.cnb[k];                     // cell = number of breakpoints <= t
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if ((int64_t)bp[mid] <= t) lo = mid + 1; else hi = mid;
        }
        c[k] = lo;
    }
    // index = (((((cT*nH + cH)*nC + cC)*nE + cE)*nW + cW)*nD + cD)
    int64_t idx = c[0];
    idx = idx * (W.cnb[1] + 1) + c[1];
    idx = idx * (W.cnb[2] + 1) + c[2];
    idx = idx * (W.cnb[3] + 1) + c[3];
    idx = idx * (W.cnb[5] + 1) + c[5];
    idx = idx * (W.cnb[4] + 1) + c[4];
    const uint32_t e = W.centry[idx];
    const int cnt = (int)(e & 0xFF);
    if (!cnt) return false;
    const uint16_t *L = W.clist + (e >> 8);
    uint64_t best = ~0ULL;
    int bl = -1, nt = 0;
    for (int q = 0; q < cnt; q++) {
        const int l = L[q];
        const uint64_t dq = pdist(W.tree[l], tg);
        if (dq < best) { best = dq; bl = l; nt = 1; }
        else if (dq == best) nt++;
    }
    f.leaf = bl;
    f.biome = W.tree[bl].biome;
    f.d = best;
    f.tied = nt >= 2;
    f.single = false;
    return true;
}

// ---------------------------------------------------------------------------------------------------
// biome of one quart cell (world_biome_cell, biome.c:87-120; mcgpu_biome.cu:54-96) as a per-lane state
// machine around one program evaluation, so the kernel has a single evaluation site whatever the lanes
// are doing (first cell, walking back along a tie chain, resolving it forward, an End chunk)
// ---------------------------------------------------------------------------------------------------
__device__ __forceinline__ int clamp_qy(const PWorld &W, int qy) {
    const int lo = W.min_qy, hi = W.min_qy + W.qh - 1;
    return qy < lo ? lo : (qy > hi ? hi : qy);
}
enum { ST_BACK = 0, ST_FWD = 1, ST_END = 2, ST_DONE = 3 };
struct Cell {
    int st;                 // ST_*
    int bxq, bzq, idx0;     // MULTI: chunk corner (quarts) and fill index of the target cell
    int k, n, leaf;         // walk position, chain length, leaf carried forward
    int ex, ey, ez;         // END: evaluation point
    int result;             // biome or -1
};
__device__ __forceinline__ void cell_start(const PWorld &W, int qx, int qy, int qz, Cell &c) {
    qy = clamp_qy(W, qy);
    c.k = 0; c.n = 0; c.leaf = -1; c.result = -1;
    if (W.source == MCG_BS_FIXED) { c.st = ST_DONE; c.result = W.fixed_biome; return; }
    if (W.source == MCG_BS_END) {
        const int cx = qx >> 2, cz = qz >> 2;
        if ((int64_t)cx * cx + (int64_t)cz * cz <= 4096LL) { c.st = ST_DONE; c.result = W.end_ids[0]; return; }
        c.st = ST_END; c.ex = (cx * 2 + 1) * 8; c.ey = qy * 4; c.ez = (cz * 2 + 1) * 8;
        return;
    }
    const int ly = qy - W.min_qy;
    c.bxq = qx & ~3; c.bzq = qz & ~3;
    c.idx0 = (ly >> 2) * 64 + (qx & 3) * 16 + (ly & 3) * 4 + (qz & 3);    // fill order of the chunk
    c.st = ST_BACK;
}
// End chunk given directly (the grid's chunk lattice)
__device__ __forceinline__ void chunk_start(const PWorld &W, int cx, int qy, int cz, Cell &c) {
    c.k = 0; c.n = 0; c.leaf = -1; c.result = -1;
    if ((int64_t)cx * cx + (int64_t)cz * cz <= 4096LL) { c.st = ST_DONE; c.result = W.end_ids[0]; return; }
    c.st = ST_END; c.ex = (cx * 2 + 1) * 8; c.ey = qy * 4; c.ez = (cz * 2 + 1) * 8;
}
__device__ __forceinline__ void cell_point(const PWorld &W, const Cell &c, int &x, int &y, int &z) {
    if (c.st == ST_END) { x = c.ex; y = c.ey; z = c.ez; return; }
    const int idx = c.idx0 - c.k, s = idx >> 6, r = idx & 63;
    x = (c.bxq + (r >> 4)) * 4;
    y = (W.min_qy + s * 4 + ((r >> 2) & 3)) * 4;
    z = (c.bzq + (r & 3)) * 4;
}
// Climate.target's quantization of the six root values (false: a value is non-finite or beyond the exact
// range of the cast; the cell is unresolved)
template <class Val>
__device__ __forceinline__ bool cell_targets(Val val, int64_t *tg) {
    bool ok = true;
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const float f = (float)val(q) * 10000.0f;
        if (!(fabsf(f) < 9.0e18f)) ok = false;     // NaN / inf / beyond the exact range of the cast
        tg[q] = ok ? (int64_t)f : 0;
    }
    return ok;
}
// advance the lane's state machine with the search result of its current cell (listed: from the cell
// table; else the tree is walked here)
__device__ __forceinline__ void cell_advance(const PWorld &W, Cell &c, const int64_t *tg, bool listed, Found f) {
    if (!listed) {
        f.leaf = rt_find(W.tree, W.ntree, tg);
        if (f.leaf < 0) { c.st = ST_DONE; c.result = -1; return; }
        f.d = pdist(W.tree[f.leaf], tg);
        f.biome = W.tree[f.leaf].biome;
        f.single = false;
    }
    const int l0 = f.leaf;
    if (c.st == ST_BACK) {
        const int idx = c.idx0 - c.k;
        if (idx == 0 || !(listed ? f.tied : rt_tie(W.tree, W.ntree, tg, f.d))) {
            c.leaf = l0;
            c.n = c.k;
            if (c.n == 0) { c.st = ST_DONE; c.result = f.biome; return; }
            c.st = ST_FWD;
            c.k = c.n - 1;                         // re-evaluate the tied cells forward
            return;
        }
        if (c.k == W.chain_cap) { c.st = ST_DONE; c.result = -1; return; }
        c.k++;
        return;
    }
    // forward: a tied cell keeps the previous cell's leaf when it is at the least distance (a cell with
    // one listed leaf has no other leaf at its least distance)
    c.leaf = !f.single && pdist(W.tree[c.leaf], tg) == f.d ? c.leaf : l0;
    if (c.k == 0) { c.st = ST_DONE; c.result = W.tree[c.leaf].biome; return; }
    c.k--;
}
// consume the program's values at the point cell_point gave; val(q) = root q's value (0 temperature,
// 1 vegetation, 2 continents, 3 erosion, 4 depth, 5 ridges)
template <class Val>
__device__ __forceinline__ void cell_step(const PWorld &W, Val val, Cell &c) {
    if (c.st == ST_END) {                          // TheEndBiomeSource (biome.c:26-45)
        const double h = (double)val(3);
        c.result = h > 0.25 ? W.end_ids[1] : (h >= -0.0625 ? W.end_ids[2] : (h < -0.21875 ? W.end_ids[3] : W.end_ids[4]));
        c.st = ST_DONE;
        return;
    }
    int64_t tg[6];
    if (!cell_targets(val, tg)) { c.st = ST_DONE; c.result = -1; return; }
    Found f;
    const bool listed = cell_search(W, tg, f);
    cell_advance(W, c, tg, listed, f);
}

__device__ __forceinline__ void count_bad(int32_t *nbad, bool bad) {
    const uint64_t m = __ballot(bad);
    if (m && nbad && (int)(threadIdx.x & 63) == __builtin_ctzll(__ballot(1))) atomicAdd(nbad, (int)__popcll(m));
}

// Quart coordinates of cell i of a router job: a chunk cell (chunk * cells + (qy * 4 + qz) * 4 + qx), a
// lattice cell (layer * nqx * nqz + column), or a cell whose coordinates the first launch stored in the
// planes cq
__device__ __forceinline__ void list_cell(const PWorld &W, const PJob &J, int64_t i, int *q) {
    if (J.geom == GEN_CHUNKS) {
        const int64_t ch = i / J.cells;
        const int r = (int)(i - ch * J.cells);
        q[0] = J.xyz[2 * ch] * 4 + (r & 3);
        q[1] = W.min_qy + (r >> 4);
        q[2] = J.xyz[2 * ch + 1] * 4 + ((r >> 2) & 3);
        return;
    }
    if (J.cq) {
        q[0] = J.cq[i]; q[1] = J.cq[J.ncells + i]; q[2] = J.cq[2 * J.ncells + i];
        return;
    }
    const int64_t per = (int64_t)J.nqx * J.nqz, ly = i >= per ? 1 : 0, col = i - ly * per;
    q[0] = J.q0x + (int)(col % J.nqx);
    q[1] = J.qys[ly];
    q[2] = J.q0z + (int)(col / J.nqx);
}

// The item loop of the cell kernels: each lane walks its items (cells of a lattice, points, grid points or
// chunk cells; stride gridDim*256) and runs one climate evaluation per loop trip, eval(c, x, y, z, same)
// (same: the lane's previous evaluation had this x and z), which also advances the lane's state machine; a
// lane whose item finishes starts its next item in the same trip.
template <class Eval>
__device__ __forceinline__ void cells_items(const PWorld &W, const PJob &J, Eval eval) {
    const int64_t stride = (int64_t)gridDim.x * kWG;
    int64_t it = (int64_t)blockIdx.x * kWG + threadIdx.x;
    Cell c;
    c.st = ST_DONE;
    c.result = 0;
    bool have = false;
    // a lattice item is a column: its two quart y layers run back to back in the lane (layer 1 reuses the
    // y-independent values of layer 0); out[layer * per + column]
    const bool lat = J.gen == GEN_LATTICE, lst = J.gen == GEN_LIST;
    const int64_t per = (int64_t)J.nqx * J.nqz, nitems = lat ? per : (lst ? (int64_t)*J.nlist : J.n);
    int layer = 0;
    int lx = 0, lz = 0;                                  // x, z of the lane's last evaluation
    bool lv = false;
    for (;;) {
        // finish the current item, start the next one (items whose biome needs no evaluation finish here)
        while (c.st == ST_DONE) {
            if (have) {
                J.out[lat ? it + layer * per : (lst ? (int64_t)J.list[it] : it)] = c.result < 0 ? (uint8_t)255 : (uint8_t)c.result;
                if (J.nbad && c.result < 0) atomicAdd(J.nbad, 1);
                if (lat && layer == 0) {
                    layer = 1;
                } else {
                    layer = 0;
                    it += stride;
                }
            }
            have = it < nitems;
            if (!have) break;
            if (lst) {
                const int64_t i = J.list[it];
                int q[3];
                list_cell(W, J, i, q);
                cell_start(W, q[0], q[1], q[2], c);
            } else if (J.gen == GEN_CHUNKS) {
                const int ch = (int)(it / J.cells), r = (int)(it % J.cells);
                cell_start(W, J.xyz[2 * ch] * 4 + (r & 3), W.min_qy + (r >> 4), J.xyz[2 * ch + 1] * 4 + ((r >> 2) & 3), c);
            } else if (lat) {
                const int r = (int)it;
                const int ax = J.q0x + r % J.nqx, az = J.q0z + r / J.nqx, qy = J.qys[layer];
                if (W.source == MCG_BS_END) chunk_start(W, ax, qy, az, c);       // a lattice of chunks
                else cell_start(W, ax, qy, az, c);
            } else {
                int x, y, z;
                if (J.gen == GEN_POINTS) {
                    x = J.xyz[3 * it]; y = J.xyz[3 * it + 1]; z = J.xyz[3 * it + 2];
                } else {
                    x = J.x0 + (int)(it % J.nx) * J.step; y = J.y; z = J.z0 + (int)(it / J.nx) * J.step;
                }
                int q[3];
                dev::zoom_quart(W.zoom_seed, x, y, z, q);
                cell_start(W, q[0], q[1], q[2], c);
            }
        }
        if (!have) break;
        int x, y, z;
        cell_point(W, c, x, y, z);
        eval(c, x, y, z, lv && x == lx && z == lz);
        lx = x; lz = z; lv = true;
    }
}

// The program interpreter over the cell items.
template <bool NEWF, bool PL, int SPD>
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(NEWF && SPD <= 6 ? 5 : 1))) void k_cells(PWorld W, PJob J) {
    typedef typename VT_of<NEWF>::T VT;
    extern __shared__ __align__(16) uint8_t lds[];
    constexpr int gbytes = grad_bytes<NEWF>();
    const int pbytes = PL ? W.noct * kPermStride + rec_bytes(W) : 0;
    dev::init_grad((typename GradOf<NEWF>::T *)lds);
    if (PL) {
        uint32_t *dst = (uint32_t *)(lds + gbytes);     // word w of octave o: table bytes 4(w & 63).., twice
        const int wpo = kPermStride / 4;                 // words per staged octave
        for (int i = threadIdx.x; i < W.noct * wpo; i += kWG) dst[i] = ((const uint32_t *)W.octs[i / wpo].p)[i & 63];
        // records: noise and layers copied word by word, octave offsets from the head of each McgOct
        uint32_t *rn = (uint32_t *)noise_of<NEWF, true>(W, lds, 0), *rl = (uint32_t *)layer_of<NEWF, true>(W, lds, 0);
        const int nn = W.nnoise * (int)sizeof(McgNoise) / 4, nl = W.nlayers * (int)sizeof(McgLayer) / 4;
        for (int i = threadIdx.x; i < nn; i += kWG) rn[i] = ((const uint32_t *)W.noise)[i];
        for (int i = threadIdx.x; i < nl; i += kWG) rl[i] = ((const uint32_t *)W.layers)[i];
        Off3 *ro = (Off3 *)off_of<NEWF, true>(W, lds, 0);
        for (int i = threadIdx.x; i < W.noct; i += kWG) ro[i] = Off3{W.octs[i].xo, W.octs[i].yo, W.octs[i].zo};
    }
    __syncthreads();
    VT *vals = (VT *)(lds + gbytes + pbytes) + threadIdx.x;
    const uint8_t *lp = lds;
    cells_items(W, J, [&](Cell &c, int x, int y, int z, bool same) {
        eval_prog<NEWF, PL, SPD>(W, lp, vals, x, y, z, same);
        cell_step(W, [&](int q) { return vals[W.root_slot[q] * kWG]; }, c);
    });
}

// The vanilla Overworld router (mcg_router.cpp) through the engine's fused column code: the seed's 46
// permutation tables doubled in LDS, the engine's gradient table and offset spline beside them, the five
// noises and the offset computed once per (x, z) and kept in registers while the lane stays on that
// column; depth per y.  MODE 1: 26.3+ (float noise), 0: 26.1/26.2 (double).  The six fields are the values
// the program's roots would hold (as floats: the search reads them as floats), so the search and the tie
// chain are the interpreter's.
// vw and vslot are W.vw and W.vslot passed again as read-only, non-aliasing kernel arguments: the
// constants and octave records are then read through the scalar cache (from the struct fields the
// compiler cannot rule out stores to them and reads them per lane).
template <int MODE, bool WRAP>
__global__ __launch_bounds__(kWG) void k_cells_van(PWorld W, PJob J, const McbWorld *__restrict__ vw,
                                                   const McbSlot *__restrict__ vslot) {
    __shared__ __attribute__((aligned(16))) uint8_t perm[MCB_OW_OCT * OCT_BYTES];
    __shared__ __attribute__((aligned(16))) typename dev::GT<MODE>::T G[256];
    __shared__ dev::SplLds spl[1];
    dev::init_grad(G);
    dev::stage_spline(spl, vw->spline);
    dev::stage_perm2(perm, vslot, MCB_OW_OCT);
    __syncthreads();
    const McbOct *oct = vslot->oct;
    dev::Noises5<MODE> n{};
    dev::Column col{};
    cells_items(W, J, [&](Cell &c, int x, int y, int z, bool same) {
        if (!same) {
            n = dev::ow_noises<MODE, const uint8_t *, WRAP>(vw, perm, G, oct, x >> 2, z >> 2);
            dev::column_finish<MODE>(vw, spl, n.t, n.h, n.c, n.e, n.w, col);
        }
        const float v[6] = {(float)n.t, (float)n.h, (float)n.c, (float)n.e, dev::ow_depth_raw<MODE>(vw, col, y), (float)n.w};
        cell_step(W, [&](int q) { return v[q]; }, c);
    });
}

// First launch, lattice: per column the climate of both quart layers, quantized as Climate.target does
// ((long)(v * 10000.0f)), as int32 in tgt[k * ncells + cell] for the tree dimensions k = temperature,
// humidity, continentalness, erosion, depth, weirdness (cell = layer * per + column).  A target that is
// non-finite or beyond +-32767 (outside the cell locator) is stored as kNoTarget: the cell goes to the
// chain windows, which apply the plugin's rules to it.
constexpr int32_t kNoTarget = INT32_MIN;
__device__ __forceinline__ int32_t target16(float v) {
    const float f = v * 10000.0f;
    return f > -32769.0f && f < 32768.0f ? (int32_t)f : kNoTarget;   // false for NaN
}
// the router's LDS: doubled permutation tables, gradient table, offset spline
// Workgroup size of the column and point launches.  Their LDS tables (31-33 KB) bound a CU to five
// workgroups: eight waves per workgroup instead of four lift the bound from 5 waves per SIMD to the
// registers' 5-7 (libmcgen's 2048^2 step-4 map 1.13 -> 1.08 ms).  The chunk launch keeps kWG: its 65,536
// lanes would make fewer workgroups than CUs.
constexpr int kVanWG = 512;
template <int MODE>
struct VanLds {
    uint8_t perm[MCB_OW_OCT * OCT_BYTES];
    typename dev::GT<MODE>::T G[256];
    dev::SplLds spl;
};
template <int MODE>
__device__ __forceinline__ void van_stage(VanLds<MODE> &L, const McbWorld *__restrict__ vw, const McbSlot *__restrict__ vslot) {
    dev::init_grad(L.G);
    dev::stage_spline(&L.spl, vw->spline);
    dev::stage_perm2(L.perm, vslot, MCB_OW_OCT);
    __syncthreads();
}
__device__ __forceinline__ void put_targets(int32_t *__restrict__ tgt, int64_t nc, int64_t cell, const int32_t *t5, int32_t dep) {
    const bool ok = t5[0] != kNoTarget && t5[1] != kNoTarget && t5[2] != kNoTarget && t5[3] != kNoTarget && t5[4] != kNoTarget;
    tgt[cell] = ok ? t5[0] : kNoTarget;
    tgt[nc + cell] = t5[1];
    tgt[2 * nc + cell] = t5[2];
    tgt[3 * nc + cell] = t5[3];
    tgt[4 * nc + cell] = dep;
    tgt[5 * nc + cell] = t5[4];
}
template <int MODE, bool WRAP>
__global__ __launch_bounds__(kVanWG) void k_van_columns(PJob J, const McbWorld *__restrict__ vw, const McbSlot *__restrict__ vslot,
                                                     int32_t *__restrict__ tgt) {
    __shared__ __attribute__((aligned(16))) VanLds<MODE> L;
    van_stage<MODE>(L, vw, vslot);
    const McbOct *oct = vslot->oct;
    const int per = J.nqx * J.nqz;
    const int64_t nc = 2 * (int64_t)per;
    for (int col = blockIdx.x * kVanWG + threadIdx.x; col < per; col += gridDim.x * kVanWG) {
        const int qx = J.q0x + col % J.nqx, qz = J.q0z + col / J.nqx;
        const dev::Noises5<MODE> n = dev::ow_noises<MODE, const uint8_t *, WRAP>(vw, L.perm, L.G, oct, qx, qz);
        dev::Column c;
        dev::column_finish<MODE>(vw, &L.spl, n.t, n.h, n.c, n.e, n.w, c);
        const int32_t t5[5] = {target16((float)n.t), target16((float)n.h), target16((float)n.c), target16((float)n.e),
                               target16((float)n.w)};
        put_targets(tgt, nc, col, t5, target16(dev::ow_depth_raw<MODE>(vw, c, J.qys[0] * 4)));
        put_targets(tgt, nc, per + col, t5, target16(dev::ow_depth_raw<MODE>(vw, c, J.qys[1] * 4)));
    }
}
// First launch, chunks (libmcgen's stored chunk biomes): a lane per chunk column (16 per chunk), its climate
// evaluated once, then the depth target of each of the world's qh quart layers; cell = chunk * cells +
// layer * 16 + (qz & 3) * 4 + (qx & 3), the chunk output layout.
template <int MODE, bool WRAP>
__global__ __launch_bounds__(kWG) void k_van_chunks(PWorld W, PJob J, const McbWorld *__restrict__ vw,
                                                    const McbSlot *__restrict__ vslot, int32_t *__restrict__ tgt) {
    __shared__ __attribute__((aligned(16))) VanLds<MODE> L;
    van_stage<MODE>(L, vw, vslot);
    const McbOct *oct = vslot->oct;
    const int64_t nc = J.n, nchunks = J.n / J.cells;
    for (int64_t it = (int64_t)blockIdx.x * kWG + threadIdx.x; it < nchunks * 16; it += (int64_t)gridDim.x * kWG) {
        const int64_t ch = it >> 4;
        const int k = (int)(it & 15);
        const int qx = J.xyz[2 * ch] * 4 + (k & 3), qz = J.xyz[2 * ch + 1] * 4 + (k >> 2);
        const dev::Noises5<MODE> n = dev::ow_noises<MODE, const uint8_t *, WRAP>(vw, L.perm, L.G, oct, qx, qz);
        dev::Column c;
        dev::column_finish<MODE>(vw, &L.spl, n.t, n.h, n.c, n.e, n.w, c);
        const int32_t t5[5] = {target16((float)n.t), target16((float)n.h), target16((float)n.c), target16((float)n.e),
                               target16((float)n.w)};
        for (int ly = 0; ly < W.qh; ly++)
            put_targets(tgt, nc, ch * J.cells + ly * 16 + k, t5, target16(dev::ow_depth_raw<MODE>(vw, c, (W.min_qy + ly) * 4)));
    }
}
// First launch, points (GEN_POINTS: blocks xyz; GEN_GRIDPTS: the grid's blocks): BiomeManager zoom to a
// quart cell, its y clamped into the world, then that cell's climate; the six targets as above and the
// cell's quart x, y, z in planes 6-8.
template <int MODE, bool WRAP>
__global__ __launch_bounds__(kVanWG) void k_van_points(PWorld W, PJob J, const McbWorld *__restrict__ vw,
                                                    const McbSlot *__restrict__ vslot, int32_t *__restrict__ tgt) {
    __shared__ __attribute__((aligned(16))) VanLds<MODE> L;
    van_stage<MODE>(L, vw, vslot);
    const McbOct *oct = vslot->oct;
    const int64_t nc = J.n;
    for (int64_t it = (int64_t)blockIdx.x * kVanWG + threadIdx.x; it < nc; it += (int64_t)gridDim.x * kVanWG) {
        int x, y, z;
        if (J.gen == GEN_POINTS) {
            x = J.xyz[3 * it]; y = J.xyz[3 * it + 1]; z = J.xyz[3 * it + 2];
        } else {
            x = J.x0 + (int)(it % J.nx) * J.step; y = J.y; z = J.z0 + (int)(it / J.nx) * J.step;
        }
        int q[3];
        dev::zoom_quart(W.zoom_seed, x, y, z, q);
        const int qy = clamp_qy(W, q[1]);
        const dev::Noises5<MODE> n = dev::ow_noises<MODE, const uint8_t *, WRAP>(vw, L.perm, L.G, oct, q[0], q[2]);
        dev::Column c;
        dev::column_finish<MODE>(vw, &L.spl, n.t, n.h, n.c, n.e, n.w, c);
        const int32_t t5[5] = {target16((float)n.t), target16((float)n.h), target16((float)n.c), target16((float)n.e),
                               target16((float)n.w)};
        put_targets(tgt, nc, it, t5, target16(dev::ow_depth_raw<MODE>(vw, c, qy * 4)));
        tgt[6 * nc + it] = q[0];
        tgt[7 * nc + it] = qy;
        tgt[8 * nc + it] = q[2];
    }
}

// Second launch: one lane per cell, the locator staged in LDS.  A cell with one listed leaf, or an untied
// least distance, or a tie at its chunk's first fill cell (which takes the first leaf), is settled here;
// the others (ties to follow along the fill order, unlisted cells, targets beyond the locator) are written
// to fix for the chain windows.  Small and memory-bound: many waves per SIMD hide the entry and record
// latencies.
__global__ __launch_bounds__(kWG) void k_van_search(PWorld W, PJob J, const int32_t *__restrict__ tgt,
                                                    int32_t *__restrict__ fix, int32_t *__restrict__ cnt) {
    __shared__ uint32_t lut[6 * 256];
    for (int i = threadIdx.x; i < 6 * 256; i += kWG) lut[i] = W.clut[i];
    __syncthreads();
    const int64_t nc = J.ncells;
    for (int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x; i < nc; i += (int64_t)gridDim.x * kWG) {
        // tree dimensions: temperature, humidity, continentalness, erosion, depth, weirdness
        const int t[6] = {tgt[i], tgt[nc + i], tgt[2 * nc + i], tgt[3 * nc + i], tgt[4 * nc + i], tgt[5 * nc + i]};
        int b = -1;
        if (t[0] != kNoTarget && t[4] != kNoTarget) {
            uint32_t idx = 0;
#pragma unroll
            for (int k = 0; k < 6; k++) {
                const uint32_t u = (uint32_t)(t[k] + 32768);
                const uint32_t e = lut[k * 256 + (int)(u >> 8)];
                idx += (e & 0xFFFFFu) + ((u & 255u) >= (e >> 20) ? (uint32_t)W.cstride[k] : 0u);
            }
            const uint32_t e = W.cent1[idx];
            const int cnt1 = (int)(e & 0xFF);
            if (cnt1 == 1) {
                b = (int)(e >> 24);
            } else if (cnt1 > 1) {
                const PRec *R = W.crec + (e >> 8);
                double best = __builtin_inf();
                int bl = 0, nt = 0;
                for (int q = 0; q < cnt1; q++) {
                    const PRec r = R[q];
                    const double dq = rdist_d(r, t);
                    if (dq < best) { best = dq; bl = r.leaf; nt = 1; }
                    else if (dq == best) nt++;
                }
                // a tie is settled here only at the chunk's first fill cell (index 0: lowest quart layer,
                // x and z multiples of 4), where the chain does not walk back
                if (nt >= 2) {
                    int q[3];
                    list_cell(W, J, i, q);
                    if (q[1] == W.min_qy && (q[0] & 3) == 0 && (q[2] & 3) == 0) nt = 1;
                }
                if (nt < 2) b = (bl >> 16) & 0xFF;
            }
        }
        if (b >= 0) {
            J.out[i] = (uint8_t)b;
        } else {
            J.out[i] = 255;
            fix[atomicAdd(cnt, 1)] = (int32_t)i;
        }
    }
}

// Third launch: the chain windows of the first kWinCells listed cells, one lane per (cell, step back j):
// the cell j places before it in its chunk's fill order evaluated and searched as the state machine's trip
// at chain position j would (targets, least distance, tie by the table or the tree).
template <int MODE, bool WRAP>
__global__ __launch_bounds__(kWG) void k_van_window(PWorld W, PJob J, const McbWorld *__restrict__ vw,
                                                    const McbSlot *__restrict__ vslot, const int32_t *__restrict__ fix,
                                                    const int32_t *__restrict__ cnt, WinCell *__restrict__ win) {
    __shared__ __attribute__((aligned(16))) VanLds<MODE> L;
    const int nf = min(cnt[0], kWinCells);
    if ((int64_t)blockIdx.x * kWG >= (int64_t)nf * kWin) return;        // the whole workgroup has nothing to do
    van_stage<MODE>(L, vw, vslot);
    const McbOct *oct = vslot->oct;
    for (int w = blockIdx.x * kWG + threadIdx.x; w < nf * kWin; w += gridDim.x * kWG) {
        const int f = w / kWin, j = w - f * kWin;
        int q[3];
        list_cell(W, J, fix[f], q);
        Cell c;
        cell_start(W, q[0], q[1], q[2], c);
        WinCell r;
        r.ok = 0;
        r.tied = r.single = 0;
        if (c.idx0 - j >= 0) {
            c.k = j;
            int x, y, z;
            cell_point(W, c, x, y, z);
            const dev::Noises5<MODE> n = dev::ow_noises<MODE, const uint8_t *, WRAP>(vw, L.perm, L.G, oct, x >> 2, z >> 2);
            dev::Column cl;
            dev::column_finish<MODE>(vw, &L.spl, n.t, n.h, n.c, n.e, n.w, cl);
            const float v[6] = {(float)n.t, (float)n.h, (float)n.c, (float)n.e, dev::ow_depth_raw<MODE>(vw, cl, y), (float)n.w};
            if (cell_targets([&](int k) { return v[k]; }, r.tg)) {
                Found fd;
                const bool listed = cell_search(W, r.tg, fd);
                if (!listed) {
                    fd.leaf = rt_find(W.tree, W.ntree, r.tg);
                    if (fd.leaf >= 0) {
                        fd.d = pdist(W.tree[fd.leaf], r.tg);
                        fd.biome = W.tree[fd.leaf].biome;
                        fd.single = false;
                        fd.tied = rt_tie(W.tree, W.ntree, r.tg, fd.d);
                    }
                }
                if (fd.leaf >= 0) {
                    r.ok = 1;
                    r.leaf = fd.leaf;
                    r.biome = fd.biome;
                    r.d = fd.d;
                    r.tied = fd.tied;
                    r.single = fd.single;
                }
            }
        }
        win[w] = r;
    }
}

// Fourth launch: each listed cell's chain from its window, as the state machine walks it (back while the
// least distance is tied and the fill index is not 0, unresolved at the cap; then forward, keeping the
// carried leaf where it is at the cell's least distance).  Cells beyond kWinCells, and chains that run past
// the window, go to the second list for the state machine.
__global__ __launch_bounds__(kWG) void k_van_chain(PWorld W, PJob J, const int32_t *__restrict__ fix, int32_t *__restrict__ cnt,
                                                   const WinCell *__restrict__ win, int32_t *__restrict__ list2) {
    const int nf = cnt[0];
    for (int f = blockIdx.x * kWG + threadIdx.x; f < nf; f += gridDim.x * kWG) {
        const int32_t i = fix[f];
        if (f >= kWinCells) { list2[atomicAdd(cnt + 1, 1)] = i; continue; }
        int q[3];
        list_cell(W, J, i, q);
        Cell c;
        cell_start(W, q[0], q[1], q[2], c);
        const WinCell *R = win + (int64_t)f * kWin;
        int n = -1, result = -2;                      // -2: past the window
        for (int k = 0; k < kWin; k++) {
            if (!R[k].ok) { result = -1; break; }
            if (c.idx0 - k == 0 || !R[k].tied) { n = k; break; }
            if (k == W.chain_cap) { result = -1; break; }
        }
        if (n < 0) {
            if (result == -2) {
                list2[atomicAdd(cnt + 1, 1)] = i;
            } else {
                J.out[i] = 255;
                if (J.nbad) atomicAdd(J.nbad, 1);
            }
            continue;
        }
        int leaf = R[n].leaf;
        if (n == 0) { J.out[i] = (uint8_t)R[0].biome; continue; }
        for (int k = n - 1; k >= 0; k--) {
            const WinCell &r = R[k];
            leaf = !r.single && pdist(W.tree[leaf], r.tg) == r.d ? leaf : r.leaf;
        }
        J.out[i] = (uint8_t)W.tree[leaf].biome;
    }
}

// the McbOct fields derived from the offsets, for the router's octaves (once per world)
__global__ void k_router_slot(McbSlot *s) {
    const int o = threadIdx.x;
    if (o < MCB_OW_OCT) s->oct[o] = dev::make_oct(s->oct[o].xo, s->oct[o].yo, s->oct[o].zo);
}

// block grid read from a lattice of cells (MULTI) or chunks (END)
// Grid zoom by tiles of 64 x zoom_rows(step) samples: the workgroup first computes the offsets of every
// lattice point its samples' corners touch (each once, into LDS), then each lane picks the corners of its
// blocks from them (the same operations in the same order as the per-block form).  A tile whose lattice
// exceeds kZoomPts (large steps) computes each block's corners directly.  Step 1 takes 32 rows and step 2
// 16 (8 and 4 blocks per lane: a tile's barriers and lattice points serve more blocks; per 4096^2 grid
// 4 rows took 142 us, 16 rows 104 us, 32 rows 97 us), larger steps 4 rows so that their lattice fits.
constexpr int kZoomTX = 64, kZoomPts = 1024;
__host__ __device__ __forceinline__ int zoom_rows(int step) { return step == 1 ? 32 : (step == 2 ? 16 : 4); }
__global__ __launch_bounds__(kWG) void k_zoom_tiles(PWorld W, PJob L, int x0, int z0, int nx, int nz, int step, int y,
                                                    uint8_t *out, int32_t *nbad) {
    __shared__ dev::ZFid F[kZoomPts];
    const int ntx = (nx + kZoomTX - 1) / kZoomTX, tz = zoom_rows(step);
    const int64_t ntiles = (int64_t)ntx * ((nz + tz - 1) / tz);
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int ix0 = (int)(t % ntx) * kZoomTX, iz0 = (int)(t / ntx) * tz;
        const int ix1 = min(ix0 + kZoomTX, nx) - 1, iz1 = min(iz0 + tz, nz) - 1;
        // lattice box of the tile: corners px..px+1 of every block, px = (x - 2) >> 2 (step > 0)
        const int pxl = (x0 + ix0 * step - 2) >> 2, pxh = ((x0 + ix1 * step - 2) >> 2) + 1;
        const int pzl = (z0 + iz0 * step - 2) >> 2, pzh = ((z0 + iz1 * step - 2) >> 2) + 1;
        const int pyl = (y - 2) >> 2;
        const int npx = pxh - pxl + 1, npz = pzh - pzl + 1;
        const bool shared = (int64_t)npx * npz * 2 <= kZoomPts;
        __syncthreads();                                  // the previous tile's readers are done
        if (shared)
            for (int i = threadIdx.x; i < npx * npz * 2; i += kWG) {
                const int u = i % npx, v = (i / npx) % npz, w = i / (npx * npz);
                F[i] = dev::zoom_fiddles(W.zoom_seed, pxl + u, pyl + w, pzl + v);
            }
        __syncthreads();
        for (int rr = 0; rr < tz; rr += kWG / kZoomTX) {          // the same trip count in every lane
        bool bad = false;
        const int ix = ix0 + (int)threadIdx.x % kZoomTX, iz = iz0 + rr + (int)threadIdx.x / kZoomTX;
        if (ix <= ix1 && iz <= iz1) {
            const int x = x0 + ix * step, z = z0 + iz * step;
            int q[3];
            if (shared)
                dev::zoom_pick(x, y, z, [&](int, int cx, int cy, int cz) {
                    return F[((cy - pyl) * npz + (cz - pzl)) * npx + (cx - pxl)];
                }, q);
            else
                dev::zoom_quart(W.zoom_seed, x, y, z, q);
            const int qy = clamp_qy(W, q[1]);
            const int layer = qy == L.qys[0] ? 0 : 1;
            // the lattice covers every cell (chunk) the zoom can pick; the range check only guards it
            const int ax = W.source == MCG_BS_END ? q[0] >> 2 : q[0], az = W.source == MCG_BS_END ? q[2] >> 2 : q[2];
            int b = 255;
            if (W.source == MCG_BS_END && (int64_t)ax * ax + (int64_t)az * az <= 4096LL) {
                b = W.end_ids[0];
            } else if (ax >= L.q0x && ax - L.q0x < L.nqx && az >= L.q0z && az - L.q0z < L.nqz) {
                b = L.out[((int64_t)layer * L.nqz + (az - L.q0z)) * L.nqx + (ax - L.q0x)];
            }
            bad = b == 255;
            out[(int64_t)iz * nx + ix] = (uint8_t)b;
        }
        count_bad(nbad, bad);
        }
    }
}

static int grid_blocks(int64_t n, int wg = kWG) {
    const int64_t b = (n + wg - 1) / wg;
    return (int)(b < 4096 ? (b > 0 ? b : 1) : 4096);
}

template <bool NEWF, bool PL>
static hipError_t launch_t(const PWorld &W, const PJob &J, int spd, int lds, hipStream_t st) {
    const int64_t items = J.gen == GEN_LATTICE ? (int64_t)J.nqx * J.nqz : J.n;     // lattice: columns
    if (W.vw) {
        hipLaunchKernelGGL((k_cells_van<NEWF ? 1 : 0, true>), dim3(grid_blocks(items)), dim3(kWG), 0, st, W, J, W.vw, W.vslot);
        return hipGetLastError();
    }
    if (spd <= 6) hipLaunchKernelGGL((k_cells<NEWF, PL, 6>), dim3(grid_blocks(items)), dim3(kWG), lds, st, W, J);
    else hipLaunchKernelGGL((k_cells<NEWF, PL, kSplineDepth>), dim3(grid_blocks(items)), dim3(kWG), lds, st, W, J);
    return hipGetLastError();
}

// End islands heights (end_height_j) of a lattice of chunks, one workgroup per tile of 16 x 16 chunks.  A
// chunk c reads section s = 2c + 1, whose window is the 25 x 25 cells around s / 2 (C division).  The tile
// tests each cell of the union of its chunks' windows once (at most 40 x 40 cells; the island test is a 2-D
// simplex) and lists the islands with their sizes in LDS; each chunk then takes Java's max over the listed
// islands inside its own window.  That max propagates NaN from either side and prefers +0 to -0 in either
// order, so the order of the list does not change the result: each value is end_height_j's.  The per-chunk
// loop evaluates up to 625 simplex cells; the tile about 6 per chunk.
template <bool OLD>
__global__ __launch_bounds__(256) void k_end_table(const McgOct *__restrict__ octs, int oct, int x0, int z0, int nx, int nz,
                                                   float *__restrict__ out) {
    constexpr int kTile = 16, kWin = kTile + 24, kCells = kWin * kWin;
    __shared__ uint8_t P[256];
    __shared__ int32_t itx[kCells], itz[kCells];
    __shared__ float isz[kCells];
    __shared__ int nisl;
    const int tiles_x = (nx + kTile - 1) / kTile;
    const int tx0 = (int)(blockIdx.x % (unsigned)tiles_x) * kTile, tz0 = (int)(blockIdx.x / (unsigned)tiles_x) * kTile;
    P[threadIdx.x] = octs[oct].p[threadIdx.x];
    if (threadIdx.x == 0) nisl = 0;
    // window centres are monotonic in the chunk, so the union of the tile's windows is a rectangle
    auto centre = [](int c) { return (2 * c + 1) / 2; };
    const int wx0 = centre(x0 + tx0) - 12, wx1 = centre(x0 + min(tx0 + kTile, nx) - 1) + 12;
    const int wz0 = centre(z0 + tz0) - 12, wz1 = centre(z0 + min(tz0 + kTile, nz) - 1) + 12;
    const int wnx = wx1 - wx0 + 1, wn = wnx * (wz1 - wz0 + 1);
    __syncthreads();
    for (int i = threadIdx.x; i < wn; i += 256) {
        const int64_t tx = wx0 + i % wnx, tz = wz0 + i / wnx;
        if (tx * tx + tz * tz <= 4096LL) continue;
        const double nv = dev::simplex2(P, (double)tx, (double)tz);
        if (OLD ? !(nv < (double)-0.9f) : !((float)nv < -0.9f)) continue;
        const int k = atomicAdd(&nisl, 1);
        itx[k] = (int)tx;
        itz[k] = (int)tz;
        isz[k] = fmodf(fabsf((float)tx) * 3439.0f + fabsf((float)tz) * 147.0f, 13.0f) + 9.0f;
    }
    __syncthreads();
    const int lx = (int)threadIdx.x % kTile, lz = (int)threadIdx.x / kTile;
    if (tx0 + lx >= nx || tz0 + lz >= nz) return;
    const int sx = 2 * (x0 + tx0 + lx) + 1, sz = 2 * (z0 + tz0 + lz) + 1;
    const int cx = sx / 2, cz = sz / 2, subx = sx % 2, subz = sz % 2;
    float doffs = -100.0f;
    if (OLD) {
        const int32_t sq = (int32_t)((uint32_t)sx * (uint32_t)sx + (uint32_t)sz * (uint32_t)sz);
        doffs = jclampf(100.0f - (float)sqrt((double)(float)sq) * 8.0f, -100.0f, 80.0f);
    }
    const int n = nisl;
    for (int k = 0; k < n; k++) {
        const int xo = itx[k] - cx, zo = itz[k] - cz;
        if (xo < -12 || xo > 12 || zo < -12 || zo > 12) continue;
        const float xd = (float)(subx - xo * 2), zd = (float)(subz - zo * 2);
        const float nd = jclampf(100.0f - (float)sqrt((double)(xd * xd + zd * zd)) * isz[k], -100.0f, 80.0f);
        doffs = jmaxf(doffs, nd);
    }
    out[(int64_t)(tz0 + lz) * nx + tx0 + lx] = doffs;
}

hipError_t launch_end_table(const PWorld &W, int oct, int x0, int z0, int nx, int nz, bool old, float *out, hipStream_t st) {
    if (nx <= 0 || nz <= 0) return hipSuccess;
    const int64_t tiles = (int64_t)((nx + 15) / 16) * ((nz + 15) / 16);
    if (tiles > INT32_MAX) return hipErrorInvalidValue;
    if (old) hipLaunchKernelGGL(k_end_table<true>, dim3((unsigned)tiles), dim3(256), 0, st, W.octs, oct, x0, z0, nx, nz, out);
    else hipLaunchKernelGGL(k_end_table<false>, dim3((unsigned)tiles), dim3(256), 0, st, W.octs, oct, x0, z0, nx, nz, out);
    return hipGetLastError();
}

hipError_t launch_cells(const PWorld &W, const PJob &J, bool newf, bool perm_lds, int lds, hipStream_t st) {
    if (J.n <= 0) return hipSuccess;
    if (newf) return perm_lds ? launch_t<true, true>(W, J, W.spl_depth, lds, st) : launch_t<true, false>(W, J, W.spl_depth, lds, st);
    return perm_lds ? launch_t<false, true>(W, J, W.spl_depth, lds, st) : launch_t<false, false>(W, J, W.spl_depth, lds, st);
}

int64_t van_cells(const PJob &J) { return J.gen == GEN_LATTICE ? 2 * (int64_t)J.nqx * J.nqz : J.n; }
int van_planes(const PJob &J) { return J.gen == GEN_LATTICE || J.gen == GEN_CHUNKS ? 6 : 9; }

hipError_t launch_van_cells(const PWorld &W, const PJob &J0, bool newf, int32_t *tgt, int32_t *fix, WinCell *win, int32_t *cnt,
                            hipStream_t st) {
    PJob J = J0;
    J.geom = J0.gen;
    J.ncells = van_cells(J0);
    if (J.ncells <= 0) return hipSuccess;
    const bool lat = J.gen == GEN_LATTICE, chk = J.gen == GEN_CHUNKS;
    J.cq = lat || chk ? nullptr : tgt + 6 * J.ncells;
    if (chk) {
        const int64_t lanes = J.n / J.cells * 16;
        if (newf) hipLaunchKernelGGL((k_van_chunks<1, true>), dim3(grid_blocks(lanes)), dim3(kWG), 0, st, W, J, W.vw, W.vslot, tgt);
        else hipLaunchKernelGGL((k_van_chunks<0, true>), dim3(grid_blocks(lanes)), dim3(kWG), 0, st, W, J, W.vw, W.vslot, tgt);
    } else if (lat) {
        const int64_t per = J.ncells / 2;
        if (newf) hipLaunchKernelGGL((k_van_columns<1, true>), dim3(grid_blocks(per, kVanWG)), dim3(kVanWG), 0, st, J, W.vw, W.vslot, tgt);
        else hipLaunchKernelGGL((k_van_columns<0, true>), dim3(grid_blocks(per, kVanWG)), dim3(kVanWG), 0, st, J, W.vw, W.vslot, tgt);
    } else {
        if (newf) hipLaunchKernelGGL((k_van_points<1, true>), dim3(grid_blocks(J.n, kVanWG)), dim3(kVanWG), 0, st, W, J, W.vw, W.vslot, tgt);
        else hipLaunchKernelGGL((k_van_points<0, true>), dim3(grid_blocks(J.n, kVanWG)), dim3(kVanWG), 0, st, W, J, W.vw, W.vslot, tgt);
    }
    if (hipError_t e = hipGetLastError()) return e;
    hipLaunchKernelGGL(k_van_search, dim3(grid_blocks(J.ncells)), dim3(kWG), 0, st, W, J, (const int32_t *)tgt, fix, cnt);
    if (hipError_t e = hipGetLastError()) return e;
    // the listed cells' count stays on the device: the window and chain launches cover the most they can
    // hold and return early past the count
    const int wblocks = grid_blocks((int64_t)kWinCells * kWin);
    if (newf)
        hipLaunchKernelGGL((k_van_window<1, true>), dim3(wblocks), dim3(kWG), 0, st, W, J, W.vw, W.vslot, (const int32_t *)fix,
                           (const int32_t *)cnt, win);
    else
        hipLaunchKernelGGL((k_van_window<0, true>), dim3(wblocks), dim3(kWG), 0, st, W, J, W.vw, W.vslot, (const int32_t *)fix,
                           (const int32_t *)cnt, win);
    if (hipError_t e = hipGetLastError()) return e;
    int32_t *list2 = fix + J.ncells;
    hipLaunchKernelGGL(k_van_chain, dim3(grid_blocks(kWinCells)), dim3(kWG), 0, st, W, J, (const int32_t *)fix, cnt,
                       (const WinCell *)win, list2);
    if (hipError_t e = hipGetLastError()) return e;
    PJob L = J;
    L.gen = GEN_LIST;
    L.list = list2;
    L.nlist = cnt + 1;
    if (newf) hipLaunchKernelGGL((k_cells_van<1, true>), dim3(64), dim3(kWG), 0, st, W, L, W.vw, W.vslot);
    else hipLaunchKernelGGL((k_cells_van<0, true>), dim3(64), dim3(kWG), 0, st, W, L, W.vw, W.vslot);
    return hipGetLastError();
}

hipError_t finish_router_slot(McbSlot *d_slot, hipStream_t st) {
    hipLaunchKernelGGL(k_router_slot, dim3(1), dim3(64), 0, st, d_slot);
    return hipGetLastError();
}

hipError_t launch_zoom_lattice(const PWorld &W, const PJob &L, int x0, int z0, int nx, int nz, int step, int y, uint8_t *out,
                               int32_t *nbad, hipStream_t st) {
    const int64_t n = (int64_t)nx * nz;
    if (n <= 0) return hipSuccess;
    const int64_t ntiles = (int64_t)((nx + kZoomTX - 1) / kZoomTX) * ((nz + zoom_rows(step) - 1) / zoom_rows(step));
    const int blocks = (int)(ntiles < 8192 ? ntiles : 8192);
    hipLaunchKernelGGL(k_zoom_tiles, dim3(blocks), dim3(kWG), 0, st, W, L, x0, z0, nx, nz, step, y, out, nbad);
    return hipGetLastError();
}

}  // namespace plug
}  // namespace mcb
