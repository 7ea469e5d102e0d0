// mcb_search.cpp - host side of the seed search (include/mcbiome.h: mcb_search, mcb_end_obs_radius).
//
// Candidates are processed in chunks: Overworld by the pruning kernel (k_search_ow, a persistent grid
// pulling candidates from a counter), everything else by seed slots (state kernels, then one wave per
// candidate over its observations).  Results are gathered once at the end; found seeds are sorted as
// unsigned like gb_search (crack/src/gpu_biomes/gpu_biomes.cu:199-240).
//
// Prefix biome sets for the pruning: for each cell of (temperature), (temperature, vegetation), ...,
// (temperature .. erosion), the union of the biomes of every leaf listed in any climate cell extending
// it.  The lists hold every possible (distance, index) minimum of their cell, so a biome missing from the
// union cannot be the result whatever the remaining targets are -- provided those targets stay inside
// the table (its outer cells reach +-2^17; the noises' bounds are checked below, else no pruning).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "mcb_search.h"

using namespace mcb;

namespace {

const uint64_t kChunkOw = 1ULL << 24;        // candidates per pruning launch
const int kChunkSlots = 32768;               // candidates per slot batch

// Quantized bound of each climate noise: |octave| <= 2 (a blend of corner dots of at most two unit
// terms), times the absolute amplitudes of both stacks (and the value factor).
bool noise_bounds_fit(const McbWorld &W) {
    for (int i = 1; i < 6; i++) {
        const McbNoise &N = W.noise[i];
        double s = 0.0;
        for (int k = 0; k < N.nlev; k++) s += W.fmt_new ? 2.0 * std::fabs((double)N.amp_f[k]) : 2.0 * std::fabs(N.amp_d[k] * N.vf_d[k]);
        if (!W.fmt_new) s *= std::fabs(N.value_factor);
        if (N.legacy || 2.0 * s * 10000.0 + 1.0 >= (double)kCellTMax) return false;
    }
    return true;
}

int ensure_feasibility(McbEngine *e) {
    if (e->d_feas) return MCB_OK;
    const CellTable &ct = *e->cells;
    int n[6];
    for (int k = 0; k < 6; k++) n[k] = ct.nb[k] + 1;
    const size_t f1 = n[0], f2 = f1 * n[1], f3 = f2 * n[2], f4 = f3 * n[3];
    std::vector<uint64_t> F4(2 * f4, 0);
    const size_t per4 = (size_t)n[5] * n[4];          // (W, depth) cells below one prefix-4 cell
    for (size_t c = 0; c < ct.entry.size(); c++) {
        const uint32_t ent = ct.entry[c];
        uint64_t m0 = 0, m1 = 0;
        if ((ent & 255u) == 0) {
            m0 = m1 = ~0ULL;                          // tree walk: any biome
        } else {
            for (uint32_t q = 0; q < (ent & 255u); q++) {
                const int b = e->tree_host[ct.list[(ent >> 8) + q]].leaf & 0xFF;
                if (b < 64) m0 |= 1ULL << b; else m1 |= 1ULL << (b - 64);
            }
        }
        const size_t p4 = c / per4;
        F4[2 * p4] |= m0;
        F4[2 * p4 + 1] |= m1;
    }
    std::vector<uint64_t> F(2 * (f1 + f2 + f3 + f4), 0);
    const size_t off[4] = {0, f1, f1 + f2, f1 + f2 + f3};
    const size_t cnt[4] = {f1, f2, f3, f4};
    for (size_t i = 0; i < 2 * f4; i++) F[2 * off[3] + i] = F4[i];
    for (int lvl = 2; lvl >= 0; lvl--) {
        const size_t child = (size_t)n[lvl + 1];
        for (size_t i = 0; i < cnt[lvl]; i++)
            for (size_t c = 0; c < child; c++) {
                F[2 * (off[lvl] + i)] |= F[2 * (off[lvl + 1] + i * child + c)];
                F[2 * (off[lvl] + i) + 1] |= F[2 * (off[lvl + 1] + i * child + c) + 1];
            }
    }
    // per dimension and bucket of 256 values: cells below | offset of the breakpoint inside << 8
    std::vector<uint32_t> lut(6 * 256);
    for (int k = 0; k < 6; k++) {
        const int32_t *bp = ct.bp.data() + ct.off[k];
        for (int b = 0; b < 256; b++) {
            const int t0 = b * 256 - 32768;
            int below = 0, off_in = 256;
            for (int i = 0; i < ct.nb[k]; i++) {
                if (bp[i] <= t0) below++;
                else if (bp[i] <= t0 + 255) off_in = bp[i] - t0;
            }
            lut[k * 256 + b] = (uint32_t)below | (uint32_t)off_in << 8;
        }
    }
    if (hipMalloc((void **)&e->d_feas, F.size() * sizeof(uint64_t)) != hipSuccess ||
        hipMalloc((void **)&e->d_coord_lut, lut.size() * sizeof(uint32_t)) != hipSuccess)
        return fail(MCB_E_NOMEM, "search tables");
    MCB_HIPCHK(hipMemcpy(e->d_feas, F.data(), F.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
    MCB_HIPCHK(hipMemcpy(e->d_coord_lut, lut.data(), lut.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    for (int i = 0; i < 4; i++) e->feas_off[i] = (int)off[i];
    e->feas_ok = e->w.host.cell_on && noise_bounds_fit(e->w.host);
    return MCB_OK;
}

}  // namespace

extern "C" {

int mcb_end_obs_radius(const McbEngine *e, const uint64_t mask[2]) {
    if (!e || !mask || e->w.dim != MCB_END) return -fail(MCB_E_ARG, "needs an End engine and a mask");
    // classes by rising height value: small_end_islands, end_barrens, end_midlands, end_highlands, and
    // the height thresholds between them (-20, 0, 40); a window cell matters only if it can raise the
    // maximum across the lowest threshold where membership changes (island size >= 9)
    const int order[4] = {3, 4, 2, 1};
    const int TH[3] = {-20, 0, 40};
    int m[4];
    for (int i = 0; i < 4; i++) {
        const int b = e->w.host.end_ids[order[i]];
        m[i] = (int)((mask[b >> 6] >> (b & 63)) & 1ULL);
    }
    int T = 1000;
    for (int i = 0; i < 3; i++)
        if (m[i] != m[i + 1] && TH[i] < T) T = TH[i];
    if (T == 1000) return 0;
    const double dmax = (101.0 - (double)T) / 9.0;
    const int rad = (int)((dmax + 1.0) / 2.0);
    return rad > 12 ? 12 : rad;
}

int mcb_search(McbEngine *e, const McbSeedSource *src, const McbObs *obs, int nobs, int flags, int64_t *found,
               uint32_t max_found, uint32_t *n_found, uint64_t *reached, double *kernel_ms) {
    if (!e || !src || !obs || nobs < 1 || nobs > MCB_MAX_OBS || !reached || (max_found && !found) ||
        src->kind < 0 || src->kind > 2 || (src->kind == 0 && src->n && !src->list))
        return fail(MCB_E_ARG, "bad arguments");
    for (int i = 0; i < nobs; i++)
        if ((obs[i].flags & 0xFF) > 12 || (obs[i].flags & ~(0xFF | MCB_OBS_BLOCK))) return fail(MCB_E_ARG, "bad observation flags");
    const uint64_t total = src->kind == 1 ? 65536ULL : src->n;
    std::lock_guard<std::mutex> lk(e->mu);
    const bool ow = e->w.dim == MCB_OVERWORLD && !(flags & MCB_SEARCH_SLOTS);
    if (ow)
        if (int rc = ensure_feasibility(e)) return rc;
    if (!e->d_obs && hipMalloc((void **)&e->d_obs, MCB_MAX_OBS * sizeof(McbObs)) != hipSuccess) return fail(MCB_E_NOMEM, "obs");
    if (!e->d_hist && hipMalloc((void **)&e->d_hist, (MCB_MAX_OBS + 1) * sizeof(unsigned long long)) != hipSuccess)
        return fail(MCB_E_NOMEM, "hist");
    if (!e->d_next && hipMalloc((void **)&e->d_next, sizeof(uint64_t)) != hipSuccess) return fail(MCB_E_NOMEM, "counter");
    if (!e->d_nfound && hipMalloc((void **)&e->d_nfound, sizeof(uint32_t)) != hipSuccess) return fail(MCB_E_NOMEM, "counter");
    if (int rc = dev_grow(&e->d_found, &e->found_cap, std::max<size_t>(1, max_found))) return rc;
    MCB_HIPCHK(hipMemcpy(e->d_obs, obs, nobs * sizeof(McbObs), hipMemcpyHostToDevice));
    MCB_HIPCHK(hipMemset(e->d_hist, 0, (nobs + 1) * sizeof(unsigned long long)));
    MCB_HIPCHK(hipMemset(e->d_nfound, 0, sizeof(uint32_t)));

    SearchArgs A;
    memset(&A, 0, sizeof A);
    A.W = e->w.d_world;
    A.tree = e->w.d_tree;
    A.cells = e->w.d_cells;
    A.crecs = e->w.d_cell_recs;
    A.feas = e->d_feas;
    for (int i = 0; i < 4; i++) A.feas_off[i] = e->feas_off[i];
    if (e->cells)
        for (int k = 0; k < 6; k++) A.ncell[k] = e->cells->nb[k] + 1;
    A.coord_lut = e->d_coord_lut;
    A.obs = e->d_obs;
    A.nobs = nobs;
    A.flags = flags | (ow && !e->feas_ok ? MCB_SEARCH_NOPRUNE : 0);
    A.src_kind = src->kind;
    A.base = src->base;
    A.next = e->d_next;
    A.found = e->d_found;
    A.nfound = e->d_nfound;
    A.max_found = max_found;
    A.hist = e->d_hist;

    int dev = 0, ncu = 0;
    MCB_HIPCHK(hipGetDevice(&dev));
    MCB_HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    hipEvent_t e0, e1;
    MCB_HIPCHK(hipEventCreate(&e0));
    MCB_HIPCHK(hipEventCreate(&e1));
    MCB_HIPCHK(hipEventRecord(e0, nullptr));
    const uint64_t chunk = ow ? kChunkOw : (uint64_t)kChunkSlots;
    if (src->kind == 0)
        if (int rc = dev_grow(&e->d_list, &e->list_cap, (size_t)std::min<uint64_t>(std::max<uint64_t>(total, 1), chunk))) return rc;
    if (!ow) {
        if (int rc = dev_grow(&e->d_slots, &e->slots_cap, (size_t)kChunkSlots)) return rc;
        if (int rc = ensure_ident(e, kChunkSlots)) return rc;
        if (int rc = dev_grow(&e->d_seeds, &e->seeds_cap, (size_t)kChunkSlots)) return rc;
    }
    for (uint64_t s0 = 0; s0 < total; s0 += chunk) {
        const uint64_t n1 = std::min(chunk, total - s0);
        A.n = n1;
        if (src->kind == 0) {
            MCB_HIPCHK(hipMemcpy(e->d_list, src->list + s0, n1 * sizeof(int64_t), hipMemcpyHostToDevice));
            A.list = e->d_list;
            A.first = 0;
        } else {
            A.first = s0;
        }
        if (ow) {
            MCB_HIPCHK(hipMemset(e->d_next, 0, sizeof(uint64_t)));
            MCB_HIPCHK(launch_search_ow(e->w, A, ncu * 2, nullptr));      // two resident workgroups per CU
        } else {
            MCB_HIPCHK(launch_seed_gen(A, (int)n1, e->d_seeds, nullptr));
            MCB_HIPCHK(launch_seed_init(e->w, e->d_seeds, e->d_ident, (int)n1, e->d_slots, nullptr));
            MCB_HIPCHK(launch_search_slots(e->w, A, e->d_slots, (int)n1, nullptr));
        }
    }
    MCB_HIPCHK(hipEventRecord(e1, nullptr));
    MCB_HIPCHK(hipEventSynchronize(e1));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (kernel_ms) *kernel_ms = ms;

    std::vector<unsigned long long> h(nobs + 1);
    MCB_HIPCHK(hipMemcpy(h.data(), e->d_hist, (nobs + 1) * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (int i = 0; i <= nobs; i++) reached[i] = h[i];
    uint32_t nf = 0;
    MCB_HIPCHK(hipMemcpy(&nf, e->d_nfound, sizeof(uint32_t), hipMemcpyDeviceToHost));
    const uint32_t keep = std::min(nf, max_found);
    std::vector<int64_t> f(keep);
    if (keep) MCB_HIPCHK(hipMemcpy(f.data(), e->d_found, keep * sizeof(int64_t), hipMemcpyDeviceToHost));
    std::sort(f.begin(), f.end(), [](int64_t a, int64_t b) { return (uint64_t)a < (uint64_t)b; });
    for (uint32_t i = 0; i < keep; i++) found[i] = f[i];
    if (n_found) *n_found = keep;
    return MCB_OK;
}

}  // extern "C"
