// mcb_search.hip - seed search: which candidate seeds show the observed biomes at the observed points
// (crack/src/gpu_biomes/gpu_biomes.cu:36-52,199-274, engine/cuda/mc_gpu.cuh:476-493,552-560,685-693).
//
// For every candidate the result is the number of observations it passes in order (the first failure
// ends it); a candidate passing all of them is found.  An observation passes when the biome at its point
// is in its set (or, with ties, when any least-distance leaf's biome is).
//
// k_search_ow<MODE> (Overworld): one lane per candidate, with no per-seed state in memory.  Each step a
// lane builds one octave's permutation in its own Fisher-Yates tile (the dominant cost, identical for
// every lane, so a wave stays converged) and adds that octave's term to its noise sums, in the
// reference's summation order.  After each climate noise the lane checks whether any biome of the
// observation's set can still win given the targets known so far (prefix cells of the climate cell
// table, see mcb_search.cpp); if none can, the candidate fails here without building the remaining
// octaves.  A lane that finishes takes the next candidate from a global counter (wave-aggregated), so
// lanes stay busy whatever stage their neighbours are in.
//
// k_search_slots<MODE, DIM> (any dimension): the candidates' full noise state is built in seed slots
// by the seed-state kernels; one wave per candidate stages it in LDS and evaluates all observations
// at once (one lane each); the number passed is the index of the first failing lane.
#include "mcb_device.h"
#include "mcb_search.h"

namespace mcb {
namespace {
using namespace dev;

__device__ __forceinline__ int64_t seed_at(const SearchArgs &A, uint64_t i) {
    if (A.src_kind == 0) return A.list[i];
    if (A.src_kind == 1) return (int64_t)(A.base | (i << 48));
    return (int64_t)(A.base + i);
}
__device__ __forceinline__ bool in_mask(const McbObs &o, int b) { return (o.mask[b >> 6] >> (b & 63)) & 1ULL; }

// Does the observation accept biome b at targets t (ties: any least-distance leaf)?
__device__ __forceinline__ bool obs_accept(const SearchArgs &A, const McbObs &o, const int *t, int b) {
    if (in_mask(o, b)) return true;
    if (!(A.flags & MCB_SEARCH_TIES)) return false;
    uint64_t m[2];
    tree_tie_mask(A.tree, A.W->tree_n, t, m);
    return ((m[0] & o.mask[0]) | (m[1] & o.mask[1])) != 0;
}

// biome at targets t = {t, h, c, e, depth, w} (cell table, tree walk beyond it)
__device__ __forceinline__ int ow_biome(const SearchArgs &A, const uint32_t *clut, const int *t) {
    const int t5[5] = {t[0], t[1], t[2], t[3], t[5]};
    uint32_t e = 0;
    if (A.W->cell_on && cell_in_range(t[0]) && cell_in_range(t[1]) && cell_in_range(t[2]) && cell_in_range(t[3]) &&
        cell_in_range(t[4]) && cell_in_range(t[5]))
        e = A.cells[cell_base(A.W, clut, t5) + cell_depth(clut, t[4])];
    if (e & 255u) return cell_best(A.crecs, e, t5, t[4]);
    int cand = -1;
    return tree_search(A.tree, A.W->tree_n, t, cand);
}

// Quart point of observation o for a seed (block observations go through BiomeManager)
__device__ __forceinline__ void obs_quart(const McbObs &o, int64_t seed, int64_t &hs, bool &have_hs, int *q) {
    if (!(o.flags & MCB_OBS_BLOCK)) {
        q[0] = o.x; q[1] = o.y; q[2] = o.z;
        return;
    }
    if (!have_hs) {
        hs = hashed_seed_dev(seed);
        have_hs = true;
    }
    zoom_quart(hs, o.x, o.y, o.z, q);
}

// ================================================================================================
// Overworld search with per-noise pruning
// ================================================================================================
// Schedule of one observation: the 6 climate noises in order (offset, temperature, vegetation,
// continentalness, erosion, ridges), each level pair in the reference's summation order.  Step s of a
// noise with n levels is (level, stack) = (s/2, s%2) for 26.3 (interleaved) and (s%n, s/n) for
// 26.1/26.2 (first stack, then second; the two stacks are separate sums).
template <int MODE>
__device__ __forceinline__ void step_level(int s, int n, int &k, int &st) {
    if (MODE) { k = s >> 1; st = s & 1; }
    else { st = s >= n; k = st ? s - n : s; }
}

// Four waves per workgroup: each lane's 256-byte Fisher-Yates tile takes 16 KB of LDS per wave, and the
// workgroup's shared tables (gradients, spline, cell locator: 14-16 KB) are staged once for four waves,
// so two workgroups fit a CU (8 waves; with two waves per workgroup only three fit: 6 waves).  The
// shuffle's LDS reads and writes are latency-bound, and the extra waves hide it: 2.25e8 -> 2.81e8
// candidates/s.  (Drawing the next index while a step's reads are in flight measured no faster.)
constexpr int kSearchWG = 256;
template <int MODE>
__global__ __launch_bounds__(kSearchWG) void k_search_ow(SearchArgs A) {
    typedef typename R<MODE>::T T_;
    __shared__ uint32_t tiles[kSearchWG / 64][64 * 64];
    __shared__ __attribute__((aligned(16))) typename GT<MODE>::T G[256];
    __shared__ SplLds spl[1];
    __shared__ uint32_t clut[MCB_CELL_LUT];
    __shared__ uint32_t hist[MCB_MAX_OBS + 1];
    const McbWorld *__restrict__ W = A.W;
    init_grad(G);
    stage_spline(spl, W->spline);
    stage_lut(clut, W);
    for (int i = threadIdx.x; i <= A.nobs; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    uint32_t *tile = tiles[threadIdx.x >> 6];
    const TileP P{reinterpret_cast<const uint8_t *>(tile) + 4 * lane};

    bool active = false, done = false;
    int64_t seed = 0, hs = 0;
    bool have_hs = false;
    Xo pos{0, 0};
    int j = 0, nn = 0, s = 0, q[3] = {0, 0, 0};
    double X = 0, Z = 0, x = 0, z = 0;
    T_ acc1 = 0, acc2 = 0, acc3 = 0, acc4 = 0;     // 26.3: acc1 (acc2 for ShiftB); 26.1: stacks 1/2 (3/4 for ShiftB)
    T_ val[6] = {0, 0, 0, 0, 0, 0};                // finished noise values (1..5)
    int qt[6] = {0, 0, 0, 0, 0, 0};                 // their quantized targets
    const bool prune = (A.flags & (MCB_SEARCH_TIES | MCB_SEARCH_NOPRUNE)) == 0;

    auto start_obs = [&]() {
        obs_quart(A.obs[j], seed, hs, have_hs, q);
        X = (double)q[0];
        Z = (double)q[2];
        nn = 0;
        s = 0;
        acc1 = acc2 = acc3 = acc4 = 0;
    };
    auto finish = [&](int passed) {
        atomicAdd(&hist[passed], 1u);
        if (passed == A.nobs) {
            const uint32_t k = atomicAdd(A.nfound, 1u);
            if (k < A.max_found) A.found[k] = seed;
        }
        active = false;
    };

    for (;;) {
        // idle lanes take the next candidates: one atomic per wave
        const bool idle = !active && !done;
        const uint64_t want = __ballot(idle);
        if (want) {
            const int leader = __ffsll((unsigned long long)want) - 1;
            uint64_t base = 0;
            if (lane == leader) base = atomicAdd((unsigned long long *)A.next, (unsigned long long)__popcll(want));
            base = __shfl(base, leader);
            if (idle) {
                const uint64_t i = base + __popcll(want & ((1ULL << lane) - 1));
                if (i < A.n) {
                    active = true;
                    seed = seed_at(A, A.first + i);
                    have_hs = false;
                    Xo root = xo_from_seed(seed);
                    pos = xo_fork(root);
                    j = 0;
                    start_obs();
                } else {
                    done = true;
                }
            }
        }
        if (!__any(active)) break;
        if (!active) continue;

        // build the octave of step s of noise nn
        const McbNoise &N = W->noise[nn];
        const int n = N.nlev;
        int k, st;
        step_level<MODE>(s, n, k, st);
        Xo r = xo_make(W->nlo[nn] ^ pos.a, W->nhi[nn] ^ pos.b);
        Xo s0 = xo_fork(r);
        Xo s1 = xo_fork(r);
        const Xo sp = st ? s1 : s0;
        Xo rk = xo_make(N.hlo[k] ^ sp.a, N.hhi[k] ^ sp.b);
        const double xo = xo_double(rk) * 256.0, yo = xo_double(rk) * 256.0, zo = xo_double(rk) * 256.0;
        fisher_yates(tile, lane, [&](int b) { return xo_bound(rk, (uint32_t)b); });
        const McbOct o = make_oct(xo, yo, zo);
        if (nn == 0) {
            const T_ ta = octave_term<MODE, false>(P, G, o, N, k, st, X, Z);
            const T_ tb = octave_term<MODE, true>(P, G, o, N, k, st, Z, X);
            if (MODE || !st) { acc1 += ta; acc2 += tb; } else { acc3 += ta; acc4 += tb; }
        } else {
            const T_ tv = octave_term<MODE, false>(P, G, o, N, k, st, x, z);
            if (MODE || !st) acc1 += tv; else acc3 += tv;
        }
        if (++s < 2 * n) continue;

        // noise nn complete
        T_ a, b = 0;
        if constexpr (MODE == 1) { a = acc1; b = acc2; }
        else { a = (acc1 + acc3) * N.value_factor; b = (acc2 + acc4) * N.value_factor; }
        acc1 = acc2 = acc3 = acc4 = 0;
        s = 0;
        if (nn == 0) {
            shift_coords<MODE>(a, b, X, Z, x, z);
            nn = 1;
            continue;
        }
        val[nn] = a;
        qt[nn] = MODE ? quantize((float)a) : quantize((float)a);
        if (prune && nn < 5) {
            // prefix cells of the targets known so far (temperature .. this noise)
            bool in = true;
            uint32_t idx = 0;
            for (int d = 1; d <= nn; d++) {
                const int dim = d - 1;                      // tree dimensions 0..3 = T, H, C, E
                in = in && cell_in_range(qt[d]);
                const uint32_t u = (uint32_t)(min(max(qt[d], -32768), 32767) + 32768);     // outer cells: clamp
                const uint32_t e = A.coord_lut[dim * 256 + (int)(u >> 8)];
                idx = idx * (uint32_t)A.ncell[dim] + (e & 0xFFu) + ((u & 255u) >= (e >> 8) ? 1u : 0u);
            }
            if (in) {
                const uint4 f = A.feas[A.feas_off[nn - 1] + idx];
                const McbObs &ob = A.obs[j];
                const uint64_t lo = (uint64_t)f.x | (uint64_t)f.y << 32, hi = (uint64_t)f.z | (uint64_t)f.w << 32;
                if (((lo & ob.mask[0]) | (hi & ob.mask[1])) == 0) {
                    finish(j);
                    continue;
                }
            }
        }
        if (nn < 5) {
            nn++;
            continue;
        }
        // all five noises: targets, biome, verdict
        Column c;
        column_finish<MODE>(W, spl, val[1], val[2], val[3], val[4], val[5], c);
        const int t[6] = {c.t, c.h, c.c, c.e, ow_depth<MODE>(W, c, q[1]), c.w};
        const int bi = ow_biome(A, clut, t);
        if (!obs_accept(A, A.obs[j], t, bi)) {
            finish(j);
            continue;
        }
        if (++j == A.nobs) {
            finish(j);
            continue;
        }
        start_obs();
    }
    __syncthreads();
    for (int i = threadIdx.x; i <= A.nobs; i += blockDim.x)
        if (hist[i]) atomicAdd(&A.hist[i], (unsigned long long)hist[i]);
}

// ================================================================================================
// Any dimension: observations of candidates whose full state is in seed slots (slot i = candidate i)
// ================================================================================================
template <int MODE, int DIM>
__global__ __launch_bounds__(64) void k_search_slots(SearchArgs A, const McbSlot *__restrict__ slots, int ncand) {
    __shared__ __attribute__((aligned(16))) uint8_t perm[(DIM == MCB_OVERWORLD ? MCB_OW_OCT : MCB_NE_OCT) * OCT_BYTES];
    __shared__ __attribute__((aligned(16))) typename GT<MODE>::T G[256];
    __shared__ uint32_t clut[DIM == MCB_OVERWORLD ? MCB_CELL_LUT : 1];
    __shared__ SplLds spl[1];
    const McbWorld *__restrict__ W = A.W;
    const int lane = threadIdx.x;
    init_grad(G);
    if (DIM == MCB_OVERWORLD) {
        stage_lut(clut, W);
        stage_spline(spl, W->spline);
    }
    const int n_oct = DIM == MCB_OVERWORLD ? MCB_OW_OCT : (DIM == MCB_NETHER ? MCB_NE_OCT : 1);
    for (int c = blockIdx.x; c < ncand; c += gridDim.x) {
        __syncthreads();
        if (DIM == MCB_END) stage_perm(perm, slots + c, 1);
        else stage_perm2(perm, slots + c, n_oct);
        __syncthreads();
        const McbOct *oct = slots[c].oct;
        const int64_t seed = slots[c].seed;
        bool pass = true;
        if (lane < A.nobs) {
            const McbObs &ob = A.obs[lane];
            int64_t hs = 0;
            bool have_hs = false;
            int q[3];
            obs_quart(ob, seed, hs, have_hs, q);
            int t[6] = {0, 0, 0, 0, 0, 0};
            int b;
            if (DIM == MCB_OVERWORLD) {
                Column col;
                ow_column<MODE>(W, perm, G, spl, oct, q[0], q[2], col);
                t[0] = col.t; t[1] = col.h; t[2] = col.c; t[3] = col.e; t[5] = col.w;
                t[4] = ow_depth<MODE>(W, col, q[1]);
                b = ow_biome(A, clut, t);
            } else if (DIM == MCB_NETHER) {
                ne_column<MODE>(W, perm, G, oct, q[0], q[2], t);
                int cand = -1;
                b = tree_search(A.tree, W->tree_n, t, cand);
            } else {
                b = end_biome<MODE>(W, perm, q[0], q[2], ob.flags & 0xFF);
            }
            pass = DIM == MCB_END ? in_mask(ob, b) : obs_accept(A, ob, t, b);
        }
        const uint64_t fail = __ballot(!pass);
        if (lane == 0) {
            const int passed = fail ? __ffsll((unsigned long long)fail) - 1 : A.nobs;
            atomicAdd(&A.hist[passed], 1ULL);
            if (passed == A.nobs) {
                const uint32_t k = atomicAdd(A.nfound, 1u);
                if (k < A.max_found) A.found[k] = seed;
            }
        }
    }
}

// candidate seeds of a chunk into a device array (slot-path input)
__global__ void k_seed_gen(SearchArgs A, int n, int64_t *out) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) out[i] = seed_at(A, A.first + (uint64_t)i);
}

}  // namespace

hipError_t launch_search_ow(const DevWorld &w, const SearchArgs &A, int blocks, hipStream_t st) {
    if (w.dim != MCB_OVERWORLD) return hipErrorInvalidValue;
    if (w.host.fmt_new) hipLaunchKernelGGL(k_search_ow<1>, dim3(blocks), dim3(kSearchWG), 0, st, A);
    else hipLaunchKernelGGL(k_search_ow<0>, dim3(blocks), dim3(kSearchWG), 0, st, A);
    return hipGetLastError();
}

hipError_t launch_search_slots(const DevWorld &w, const SearchArgs &A, const McbSlot *slots, int ncand, hipStream_t st) {
    if (ncand <= 0) return hipSuccess;
    const int blocks = ncand < 65536 ? ncand : 65536;
    const bool f = w.host.fmt_new != 0;
    switch (w.dim) {
        case MCB_OVERWORLD:
            if (f) hipLaunchKernelGGL((k_search_slots<1, MCB_OVERWORLD>), dim3(blocks), dim3(64), 0, st, A, slots, ncand);
            else hipLaunchKernelGGL((k_search_slots<0, MCB_OVERWORLD>), dim3(blocks), dim3(64), 0, st, A, slots, ncand);
            break;
        case MCB_NETHER:
            if (f) hipLaunchKernelGGL((k_search_slots<1, MCB_NETHER>), dim3(blocks), dim3(64), 0, st, A, slots, ncand);
            else hipLaunchKernelGGL((k_search_slots<0, MCB_NETHER>), dim3(blocks), dim3(64), 0, st, A, slots, ncand);
            break;
        default:
            if (f) hipLaunchKernelGGL((k_search_slots<1, MCB_END>), dim3(blocks), dim3(64), 0, st, A, slots, ncand);
            else hipLaunchKernelGGL((k_search_slots<0, MCB_END>), dim3(blocks), dim3(64), 0, st, A, slots, ncand);
    }
    return hipGetLastError();
}

hipError_t launch_seed_gen(const SearchArgs &A, int n, int64_t *out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_seed_gen, dim3((n + 255) / 256), dim3(256), 0, st, A, n, out);
    return hipGetLastError();
}

}  // namespace mcb
