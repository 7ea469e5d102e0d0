// mcb_serve.cpp - serving loop: continuous batching of region requests on one device.
//
// Requests (McbRegion of any world of a context) are queued by mcb_server_submit.  Each of `lanes`
// in-flight batch slots has a thread that forms a batch whenever its lane is free: it admits queued requests in
// arrival order while their output and zoom-staging bytes fit the lane's device arena (memory-based
// admission), looks their seeds up in a per-world device LRU of seed-state slots (only misses are
// initialised), lowers the batch into kernel descriptors and launches it on the lane's stream
// (seed state for misses, persistent quart-grid launch, zoom pass, copy-out).  Batches on different
// lanes overlap on the device; the lane's thread then waits for its batch and hands each
// request's copy-out to a pool of copy threads, which move the bytes to the callers' buffers and record
// per-request latency (submit -> result in the caller's host buffer) as each request's bytes land, so a
// request does not wait for the copies of the requests batched before it.
//
// The nearest precedent in the reference is the pinned-slot terrain producer of
// libmcgen/src/gpu_bridge.c:605-760; the reference has no request server for biomes.
#include <algorithm>
#include <array>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <memory>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "mcb_internal.h"

using namespace mcb;

struct McbContext;    // defined in mcb_plan.cpp
extern "C" int mcb_context_world_engine(McbContext *c, int world, McbEngine **e);

namespace {
using Clock = std::chrono::steady_clock;
constexpr int kCopyThreads = 8;           // threads moving results from the lanes' pinned buffers to callers

struct Req {
    McbRegion r;
    uint8_t *host_out;
    int64_t ticket;
    Clock::time_point t_submit, t_admit;
    int64_t out_bytes, stage_bytes;
    // set at admission: the world's slot pool, the seed's pinned slot, whether this batch initialises it
    struct SlotPool *pool = nullptr;
    int slot = -1;
    bool fresh = false;
};

// Device LRU of seed-state slots for one world.
struct SlotPool {
    McbEngine *e = nullptr;
    McbSlot *d_slots = nullptr;
    int cap = 0;
    std::unordered_map<int64_t, int> slot_of;
    std::vector<int64_t> seed_of;
    std::vector<int> pins;
    std::vector<int> init_lane;                           // lane whose batch initialised the slot
    std::vector<int64_t> init_batch;
    std::list<int> lru;                                   // front = most recent
    std::vector<std::list<int>::iterator> pos;
    int npinned = 0;                                      // slots with pins > 0
    int64_t hits = 0, misses = 0;

    int init(McbEngine *eng, int capacity) {
        e = eng;
        cap = capacity;
        if (hipMalloc((void **)&d_slots, (size_t)cap * sizeof(McbSlot)) != hipSuccess) return fail(MCB_E_NOMEM, "slot pool");
        seed_of.assign(cap, 0);
        pins.assign(cap, 0);
        init_lane.assign(cap, -1);
        init_batch.assign(cap, -1);
        pos.resize(cap);
        for (int i = 0; i < cap; i++) { lru.push_back(i); pos[i] = std::prev(lru.end()); seed_of[i] = INT64_MIN; }
        return 0;
    }
    // slot for seed; *fresh = 1 when it must be initialised.  -1 when every slot is pinned.
    int acquire(int64_t seed, int *fresh) {
        auto it = slot_of.find(seed);
        if (it != slot_of.end() && seed_of[it->second] == seed) {
            int s = it->second;
            lru.splice(lru.begin(), lru, pos[s]);
            if (pins[s]++ == 0) npinned++;
            hits++;
            *fresh = 0;
            return s;
        }
        for (auto r = lru.rbegin(); r != lru.rend(); ++r) {
            int s = *r;
            if (pins[s]) continue;
            if (seed_of[s] != INT64_MIN) slot_of.erase(seed_of[s]);
            seed_of[s] = seed;
            slot_of[seed] = s;
            lru.splice(lru.begin(), lru, pos[s]);
            if (pins[s]++ == 0) npinned++;
            misses++;
            *fresh = 1;
            return s;
        }
        return -1;
    }
    void release(int s) {
        if (pins[s] > 0 && --pins[s] == 0) npinned--;
    }
    // forget a slot's seed: its state was never written (the batch that acquired it failed)
    void unmap(int s) {
        auto it = slot_of.find(seed_of[s]);
        if (it != slot_of.end() && it->second == s) slot_of.erase(it);
        seed_of[s] = INT64_MIN;
        init_lane[s] = -1;
        init_batch[s] = -1;
    }
    // true when `seed` holds a pinned slot (a batch in flight uses it)
    bool pinned_seed(int64_t seed) const {
        auto it = slot_of.find(seed);
        return it != slot_of.end() && seed_of[it->second] == seed && pins[it->second] > 0;
    }
};

struct Lane {
    hipStream_t st = nullptr;
    hipEvent_t done = nullptr;
    hipEvent_t inited = nullptr;                          // after this lane's seed-state kernels
    int64_t batch_id = 0;
    uint8_t *d_out = nullptr, *d_quart = nullptr, *h_out = nullptr;
    char *d_desc = nullptr, *h_desc = nullptr;
    int64_t desc_cap = 0;
    std::vector<Req> reqs;
    std::vector<std::pair<int, int>> pinned;              // (world, slot)
    std::vector<std::pair<int, int>> fresh;               // (world, slot) acquired unmapped by this batch
    std::vector<std::pair<int, int64_t>> wait_lanes;      // (lane, batch) whose seed-state kernels this batch reads
    int64_t inited_batch = 0;                             // last batch whose seed-state event is recorded
    bool busy = false;
    Clock::time_point t_launched;                         // launch_batch returned
    Clock::time_point t_done;                             // the batch's results reached pinned memory
    int copies_left = 0;                                  // requests of the batch not yet copied out
    std::thread worker;
};
struct CopyJob {
    int lane;
    size_t req;
};
}  // namespace

struct McbServer {
    McbContext *ctx = nullptr;
    int64_t batch_bytes = 0;
    int tile_w = 16, cpl = 2;
    std::vector<Lane> lanes;
    std::unordered_map<int, SlotPool> pools;
    int slot_capacity = 0;
    std::mutex mu;
    std::condition_variable cv_submit, cv_done, cv_copy, cv_copied, cv_inited;
    std::deque<CopyJob> copies;
    std::vector<std::thread> copiers;
    std::deque<Req> pending;
    int inflight = 0;                                      // lanes with a launched batch
    int64_t completions = 0;                               // batches finished (slots unpinned)
    std::unordered_map<int64_t, std::pair<int, double>> results;   // ticket -> (status, latency ms)
    std::vector<double> latencies;
    std::vector<std::array<double, 4>> phases;             // per latency: queue, launch, device, copy (ms)
    int64_t next_ticket = 1;
    bool stop = false;
    std::string fatal;
    int64_t batches = 0, admitted = 0;
};

namespace {

int64_t region_out_bytes(const McbRegion &r) { return (int64_t)r.ny * r.nx * r.nz; }
int64_t region_stage_bytes(const McbRegion &r) {
    if (!r.zoom) return 0;
    const int64_t qx = (((r.x0 + r.nx - 1 - 2) >> 2) + 1) - ((r.x0 - 2) >> 2) + 1;
    const int64_t qz = (((r.z0 + r.nz - 1 - 2) >> 2) + 1) - ((r.z0 - 2) >> 2) + 1;
    return qx * qz * 2 * r.ny;
}

// launch one batch on lane L; called with the server lock NOT held
int launch_batch(McbServer *S, Lane &L) {
    // group by world
    std::vector<int> worlds;
    for (auto &q : L.reqs)
        if (std::find(worlds.begin(), worlds.end(), q.r.world) == worlds.end()) worlds.push_back(q.r.world);
    int64_t out_off = 0, quart_off = 0;
    std::vector<int64_t> req_out(L.reqs.size());
    // descriptor staging layout per world: seeds, slot ids, greqs, gtiles, qreqs, qtiles, zreqs, ztiles
    struct Seg { int64_t off, n; };
    struct WorldBatch {
        int world;
        SlotPool *pool;
        std::vector<int64_t> init_seeds;
        std::vector<int32_t> init_slots;
        Lowered L;
        Seg seeds, slots, g, gt, q, qt, z, zt;
    };
    std::vector<WorldBatch> wb(worlds.size());
    for (size_t wi = 0; wi < worlds.size(); wi++) {
        WorldBatch &B = wb[wi];
        B.world = worlds[wi];
        for (size_t i = 0; i < L.reqs.size(); i++) {
            Req &q = L.reqs[i];
            if (q.r.world != B.world) continue;
            B.pool = q.pool;
            const int slot = q.slot;
            if (q.fresh) {
                B.init_seeds.push_back(q.r.seed);
                B.init_slots.push_back(slot);
            }
            McbRegion r = q.r;
            r.out_off = out_off;
            req_out[i] = out_off;
            out_off += q.out_bytes;
            const size_t g0 = B.L.g.size(), q0 = B.L.q.size(), z0 = B.L.z.size();
            Lowered one;
            one.quart_bytes = quart_off;          // zoom staging offsets run across the whole lane
            lower_region(r, slot, S->tile_w, S->cpl, one);
            quart_off = one.quart_bytes;
            for (auto t : one.gt) { t.req += (int)g0; B.L.gt.push_back(t); }
            for (auto t : one.qt) { t.req += (int)q0; B.L.qt.push_back(t); }
            for (auto t : one.zt) { t.req += (int)z0; B.L.zt.push_back(t); }
            B.L.max_ny_g = std::max(B.L.max_ny_g, one.max_ny_g);
            B.L.max_ny_q = std::max(B.L.max_ny_q, one.max_ny_q);
            B.L.g.insert(B.L.g.end(), one.g.begin(), one.g.end());
            B.L.q.insert(B.L.q.end(), one.q.begin(), one.q.end());
            B.L.z.insert(B.L.z.end(), one.z.begin(), one.z.end());
        }
    }
    // order after the seed-state kernels of batches on other lanes that initialise slots this batch reads:
    // their lane records the event in its own launch, which may still be running
    for (const auto &w8 : L.wait_lanes) {
        std::unique_lock<std::mutex> lk(S->mu);
        S->cv_inited.wait(lk, [&] { return S->lanes[w8.first].inited_batch >= w8.second; });
        MCB_HIPCHK(hipStreamWaitEvent(L.st, S->lanes[w8.first].inited, 0));
    }
    // pack descriptors into the lane's pinned staging, one H2D copy
    int64_t off = 0;
    auto put = [&](Seg &s, const void *src, int64_t bytes, int64_t n) {
        off = (off + 15) & ~(int64_t)15;
        s.off = off;
        s.n = n;
        if (off + bytes <= L.desc_cap && bytes) memcpy(L.h_desc + off, src, (size_t)bytes);
        off += bytes;
    };
    for (auto &B : wb) {
        put(B.seeds, B.init_seeds.data(), (int64_t)B.init_seeds.size() * 8, (int64_t)B.init_seeds.size());
        put(B.slots, B.init_slots.data(), (int64_t)B.init_slots.size() * 4, (int64_t)B.init_slots.size());
        put(B.g, B.L.g.data(), (int64_t)B.L.g.size() * sizeof(McbRequest), (int64_t)B.L.g.size());
        put(B.gt, B.L.gt.data(), (int64_t)B.L.gt.size() * sizeof(McbTile), (int64_t)B.L.gt.size());
        put(B.q, B.L.q.data(), (int64_t)B.L.q.size() * sizeof(McbRequest), (int64_t)B.L.q.size());
        put(B.qt, B.L.qt.data(), (int64_t)B.L.qt.size() * sizeof(McbTile), (int64_t)B.L.qt.size());
        put(B.z, B.L.z.data(), (int64_t)B.L.z.size() * sizeof(McbRequest), (int64_t)B.L.z.size());
        put(B.zt, B.L.zt.data(), (int64_t)B.L.zt.size() * sizeof(McbTile), (int64_t)B.L.zt.size());
    }
    if (off > L.desc_cap) return fail(MCB_E_LIMIT, "batch descriptors exceed %lld bytes", (long long)L.desc_cap);
    MCB_HIPCHK(hipMemcpyAsync(L.d_desc, L.h_desc, (size_t)off, hipMemcpyHostToDevice, L.st));
    auto dp = [&](const Seg &s) { return (void *)(L.d_desc + s.off); };
    for (auto &B : wb)
        MCB_HIPCHK(launch_seed_init(B.pool->e->w, (const int64_t *)dp(B.seeds), (const int32_t *)dp(B.slots), (int)B.seeds.n,
                                    B.pool->d_slots, L.st));
    MCB_HIPCHK(hipEventRecord(L.inited, L.st));
    {
        std::lock_guard<std::mutex> lk(S->mu);
        L.inited_batch = L.batch_id;
        S->cv_inited.notify_all();
    }
    for (auto &B : wb) {
        const DevWorld &w = B.pool->e->w;
        MCB_HIPCHK(launch_grid(w, B.pool->d_slots, (const McbRequest *)dp(B.g), (const McbTile *)dp(B.gt), (int)B.gt.n,
                               L.d_out, S->tile_w, S->cpl, B.L.max_ny_g, (int)B.gt.n, wrap_free(w.host, B.L.g), L.st));
        MCB_HIPCHK(launch_grid(w, B.pool->d_slots, (const McbRequest *)dp(B.q), (const McbTile *)dp(B.qt), (int)B.qt.n,
                               L.d_quart, S->tile_w, S->cpl, B.L.max_ny_q, (int)B.qt.n, wrap_free(w.host, B.L.q), L.st));
        MCB_HIPCHK(launch_zoom((const McbRequest *)dp(B.z), (const McbRequest *)dp(B.q), (const McbTile *)dp(B.zt),
                               (int)B.zt.n, L.d_quart, L.d_out, L.st));
    }
    if (out_off) MCB_HIPCHK(hipMemcpyAsync(L.h_out, L.d_out, (size_t)out_off, hipMemcpyDeviceToHost, L.st));
    MCB_HIPCHK(hipEventRecord(L.done, L.st));
    for (size_t i = 0; i < L.reqs.size(); i++) L.reqs[i].r.out_off = req_out[i];
    return 0;
}

// A failed batch leaves no seed mapped to a slot whose state it never wrote, and drops its pins.
void abandon_batch(McbServer *S, Lane &L) {
    for (auto &p : L.fresh) S->pools[p.first].unmap(p.second);
    for (auto &p : L.pinned) S->pools[p.first].release(p.second);
    L.fresh.clear();
    L.pinned.clear();
}

// Admission for lane L, with the server lock held: requests in arrival order while the lane's arena holds
// their outputs and zoom staging (memory) and the world's unpinned seed slots hold the batch's new distinct
// seeds.  Empty when the next request's seed finds every slot pinned by batches in flight.
void admit(McbServer *S, Lane &L, int li) {
    L.reqs.clear();
    L.pinned.clear();
    L.fresh.clear();
    L.wait_lanes.clear();
    L.batch_id = S->batches + 1;
    int64_t out = 0, stage = 0;
    // per world: the batch's distinct seeds, and how many of them need a slot that no batch in flight has
    // pinned (an unmapped seed evicts one; a resident unpinned one pins its own)
    std::unordered_map<int, std::unordered_set<int64_t>> seen;
    std::unordered_map<int, int> need;
    while (!S->pending.empty()) {
        Req &q = S->pending.front();
        if (!L.reqs.empty() && (out + q.out_bytes > S->batch_bytes || stage + q.stage_bytes > S->batch_bytes)) break;
        const int w = q.r.world;
        const bool new_seed = seen[w].count(q.r.seed) == 0;
        bool takes_slot = false;
        if (new_seed) {
            auto pit = S->pools.find(w);
            const SlotPool *P = (pit != S->pools.end() && pit->second.e) ? &pit->second : nullptr;
            const int unpinned = P ? P->cap - P->npinned : S->slot_capacity;
            takes_slot = !(P && P->pinned_seed(q.r.seed));
            if (takes_slot && need[w] + 1 > unpinned) break;
        }
        if (new_seed) {
            seen[w].insert(q.r.seed);
            if (takes_slot) need[w]++;
        }
        // the world's slot pool (created with the world's engine on its first request)
        SlotPool &pool = S->pools[w];
        if (!pool.e) {
            McbEngine *e = nullptr;
            int rc = mcb_context_world_engine(S->ctx, w, &e);
            if (!rc) rc = pool.init(e, S->slot_capacity);
            if (rc) {
                S->results[q.ticket] = {rc, 0.0};
                S->fatal = mcb_last_error();
                S->pending.pop_front();
                S->cv_done.notify_all();
                continue;
            }
        }
        // pin the seed's slot now, so that other lanes' admissions see it taken
        int fresh = 0;
        const int slot = pool.acquire(q.r.seed, &fresh);
        if (slot < 0) break;                              // every slot pinned (admission above prevents this)
        L.pinned.push_back({w, slot});
        if (fresh) {
            L.fresh.push_back({w, slot});
            pool.init_lane[slot] = li;
            pool.init_batch[slot] = L.batch_id;
        } else {
            // state written by a batch still in flight on another lane: order after its seed-state kernels
            const int il = pool.init_lane[slot];
            const std::pair<int, int64_t> w8{il, pool.init_batch[slot]};
            if (il >= 0 && il != li && S->lanes[il].busy && S->lanes[il].batch_id == w8.second &&
                std::find(L.wait_lanes.begin(), L.wait_lanes.end(), w8) == L.wait_lanes.end())
                L.wait_lanes.push_back(w8);
        }
        q.pool = &pool;
        q.slot = slot;
        q.fresh = fresh != 0;
        out += q.out_bytes;
        stage += q.stage_bytes;
        q.t_admit = Clock::now();
        L.reqs.push_back(q);
        S->pending.pop_front();
    }
}

// Thread of lane li: forms a batch whenever requests are queued and the lane is free, launches it,
// waits for it, hands each request's copy-out to the copy threads and frees the lane once every
// request's bytes have left its pinned buffer.  One thread per lane, so a batch whose host-side launch
// stalls holds up only its own requests; the others keep admitting.
void lane_main(McbServer *S, int li) {
    Lane &L = S->lanes[li];
    int64_t blocked_at = -1;     // completions count when admission last found no free seed slot
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_submit.wait(lk, [&] {
                if (S->stop) return true;
                if (S->pending.empty()) return false;
                return !(blocked_at >= 0 && S->completions == blocked_at);
            });
            if (S->stop) return;
            blocked_at = -1;
            admit(S, L, li);
            if (L.reqs.empty()) {
                blocked_at = S->completions;
                continue;
            }
            L.busy = true;
            ++S->batches;
            S->admitted += (int64_t)L.reqs.size();
            S->inflight++;
        }
        const int rc = launch_batch(S, L);
        if (rc) {
            (void)hipStreamSynchronize(L.st);
            std::lock_guard<std::mutex> lk(S->mu);
            L.inited_batch = std::max(L.inited_batch, L.batch_id);     // waiters on this batch proceed (fatal state)
            S->cv_inited.notify_all();
            S->fatal = mcb_last_error();
            for (auto &q : L.reqs) S->results[q.ticket] = {rc, 0.0};
            abandon_batch(S, L);
            S->inflight--;
            S->completions++;
            L.busy = false;
            S->cv_done.notify_all();
            S->cv_submit.notify_all();
            continue;
        }
        L.t_launched = Clock::now();
        // poll rather than block: a blocking wait can sleep past the event by milliseconds
        hipError_t e;
        while ((e = hipEventQuery(L.done)) == hipErrorNotReady) std::this_thread::yield();
        std::unique_lock<std::mutex> lk(S->mu);
        L.t_done = Clock::now();
        if (e == hipSuccess && !L.reqs.empty()) {
            L.copies_left = (int)L.reqs.size();
            for (size_t i = 0; i < L.reqs.size(); i++) S->copies.push_back({li, i});
            S->cv_copy.notify_all();
            S->cv_copied.wait(lk, [&] { return L.copies_left == 0; });
        }
        S->inflight--;
        if (e != hipSuccess) {
            for (auto &q : L.reqs) S->results[q.ticket] = {(int)MCB_E_DEVICE, 0.0};
            S->fatal = hipGetErrorString(e);
            abandon_batch(S, L);
        } else {
            for (auto &p : L.pinned) S->pools[p.first].release(p.second);
        }
        S->completions++;
        L.busy = false;
        S->cv_done.notify_all();
        S->cv_submit.notify_all();
    }
}

// copy thread: one request's bytes from its lane's pinned buffer to the caller, then its latency
void copier_main(McbServer *S) {
    auto span = [](Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
    for (;;) {
        CopyJob j;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_copy.wait(lk, [&] { return S->stop || !S->copies.empty(); });
            if (S->copies.empty()) return;
            j = S->copies.front();
            S->copies.pop_front();
        }
        Lane &L = S->lanes[j.lane];
        const Req &q = L.reqs[j.req];
        memcpy(q.host_out, L.h_out + q.r.out_off, (size_t)q.out_bytes);
        const Clock::time_point t = Clock::now();
        const double ms = span(q.t_submit, t);
        const std::array<double, 4> ph = {span(q.t_submit, q.t_admit), span(q.t_admit, L.t_launched),
                                          span(L.t_launched, L.t_done), span(L.t_done, t)};
        std::lock_guard<std::mutex> lk(S->mu);
        S->results[q.ticket] = {0, ms};
        S->latencies.push_back(ms);
        S->phases.push_back(ph);
        if (--L.copies_left == 0) S->cv_copied.notify_all();
        S->cv_done.notify_all();
    }
}

}  // namespace

extern "C" {

int mcb_server_create(McbServer **out, McbContext *ctx, int64_t batch_bytes, int lanes, int slot_capacity) {
    if (!out || !ctx) return fail(MCB_E_ARG, "bad arguments");
    *out = nullptr;
    if (batch_bytes <= 0) batch_bytes = 8ll << 20;
    if (lanes <= 0) lanes = 4;
    if (lanes > 64) return fail(MCB_E_ARG, "at most 64 lanes");
    if (slot_capacity <= 0) slot_capacity = 16384;
    std::unique_ptr<McbServer> S(new McbServer());
    S->ctx = ctx;
    S->batch_bytes = batch_bytes;
    S->slot_capacity = slot_capacity;
    S->lanes.resize(lanes);
    for (auto &L : S->lanes) {
        // descriptors: generous bound for requests and 16x16 tiles of a full arena
        L.desc_cap = (batch_bytes / 256 + 4096) * (int64_t)sizeof(McbTile) * 2 + (64ll << 20);
        MCB_HIPCHK(hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking));
        MCB_HIPCHK(hipEventCreateWithFlags(&L.done, hipEventDisableTiming));
        MCB_HIPCHK(hipEventCreateWithFlags(&L.inited, hipEventDisableTiming));
        MCB_HIPCHK(hipMalloc((void **)&L.d_out, (size_t)batch_bytes));
        MCB_HIPCHK(hipMalloc((void **)&L.d_quart, (size_t)batch_bytes));
        MCB_HIPCHK(hipMalloc((void **)&L.d_desc, (size_t)L.desc_cap));
        MCB_HIPCHK(hipHostMalloc((void **)&L.h_out, (size_t)batch_bytes, hipHostMallocDefault));
        MCB_HIPCHK(hipHostMalloc((void **)&L.h_desc, (size_t)L.desc_cap, hipHostMallocDefault));
        // the runtime sets up a stream's queues on its first commands: a fill and both copy directions now,
        // so that no request pays for them
        MCB_HIPCHK(hipMemsetAsync(L.d_out, 0, 256, L.st));
        MCB_HIPCHK(hipMemcpyAsync(L.d_desc, L.h_desc, 256, hipMemcpyHostToDevice, L.st));
        MCB_HIPCHK(hipMemcpyAsync(L.h_out, L.d_out, 256, hipMemcpyDeviceToHost, L.st));
        MCB_HIPCHK(hipStreamSynchronize(L.st));
    }
    McbServer *s = S.release();
    for (int i = 0; i < lanes; i++) s->lanes[i].worker = std::thread(lane_main, s, i);
    for (int i = 0; i < kCopyThreads; i++) s->copiers.emplace_back(copier_main, s);
    *out = s;
    return MCB_OK;
}

int mcb_server_submit(McbServer *S, const McbRegion *r, uint8_t *host_out, int64_t *ticket) {
    if (!S || !r || !host_out) return fail(MCB_E_ARG, "bad arguments");
    if (int rc = check_region(*r)) return rc;
    Req q;
    q.r = *r;
    q.host_out = host_out;
    q.out_bytes = region_out_bytes(*r);
    q.stage_bytes = region_stage_bytes(*r);
    if (q.out_bytes > S->batch_bytes || q.stage_bytes > S->batch_bytes) return fail(MCB_E_LIMIT, "request larger than a batch arena");
    q.t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        q.ticket = S->next_ticket++;
        S->pending.push_back(q);
    }
    if (ticket) *ticket = q.ticket;
    S->cv_submit.notify_one();
    return MCB_OK;
}

int mcb_server_wait(McbServer *S, int64_t ticket, double *latency_ms) {
    if (!S) return fail(MCB_E_ARG, "null server");
    std::unique_lock<std::mutex> lk(S->mu);
    S->cv_done.wait(lk, [&] { return S->results.count(ticket) > 0; });
    auto r = S->results[ticket];
    S->results.erase(ticket);
    if (latency_ms) *latency_ms = r.second;
    if (r.first) return fail(r.first, "request failed: %s", S->fatal.c_str());
    return MCB_OK;
}

int mcb_server_drain(McbServer *S) {
    if (!S) return fail(MCB_E_ARG, "null server");
    std::unique_lock<std::mutex> lk(S->mu);
    S->cv_done.wait(lk, [&] {
        if (!S->pending.empty() || S->inflight) return false;
        for (auto &L : S->lanes) if (L.busy) return false;
        return true;
    });
    return MCB_OK;
}

static double pct(std::vector<double> v, double p) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    size_t k = (size_t)std::min<double>((double)v.size() - 1, std::max(0.0, p / 100.0 * (double)v.size() - 1.0 + 0.999999));
    return v[k];
}

int mcb_server_stats(McbServer *S, double *p50, double *p99, double *p999, int64_t *completed, double *hit_rate,
                     int64_t *batches) {
    if (!S) return fail(MCB_E_ARG, "null server");
    std::lock_guard<std::mutex> lk(S->mu);
    if (p50) *p50 = pct(S->latencies, 50.0);
    if (p99) *p99 = pct(S->latencies, 99.0);
    if (p999) *p999 = pct(S->latencies, 99.9);
    if (completed) *completed = (int64_t)S->latencies.size();
    int64_t h = 0, m = 0;
    for (auto &kv : S->pools) { h += kv.second.hits; m += kv.second.misses; }
    if (hit_rate) *hit_rate = (h + m) ? (double)h / (double)(h + m) : 0.0;
    if (batches) *batches = S->batches;
    return MCB_OK;
}

int mcb_server_slowest(McbServer *S, int n, double *out) {
    if (!S || n < 0 || (n && !out)) return fail(MCB_E_ARG, "bad arguments");
    std::lock_guard<std::mutex> lk(S->mu);
    std::vector<size_t> idx(S->latencies.size());
    for (size_t i = 0; i < idx.size(); i++) idx[i] = i;
    const size_t k = std::min<size_t>((size_t)n, idx.size());
    std::partial_sort(idx.begin(), idx.begin() + k, idx.end(), [&](size_t a, size_t b) { return S->latencies[a] > S->latencies[b]; });
    for (size_t i = 0; i < (size_t)n; i++) {
        double *o = out + 5 * i;
        if (i >= k) { for (int j = 0; j < 5; j++) o[j] = 0.0; continue; }
        o[0] = S->latencies[idx[i]];
        for (int j = 0; j < 4; j++) o[1 + j] = S->phases[idx[i]][j];
    }
    return (int)k;
}

int mcb_server_reset_stats(McbServer *S) {
    if (!S) return fail(MCB_E_ARG, "null server");
    std::lock_guard<std::mutex> lk(S->mu);
    S->latencies.clear();
    S->phases.clear();
    for (auto &kv : S->pools) { kv.second.hits = 0; kv.second.misses = 0; }
    S->batches = 0;
    return MCB_OK;
}

void mcb_server_destroy(McbServer *S) {
    if (!S) return;
    mcb_server_drain(S);
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_submit.notify_all();
    S->cv_copy.notify_all();
    for (auto &L : S->lanes) L.worker.join();
    for (auto &t : S->copiers) t.join();
    for (auto &L : S->lanes) {
        (void)hipStreamSynchronize(L.st);
        (void)hipStreamDestroy(L.st);
        (void)hipEventDestroy(L.done);
        (void)hipEventDestroy(L.inited);
        (void)hipFree(L.d_out);
        (void)hipFree(L.d_quart);
        (void)hipFree(L.d_desc);
        (void)hipHostFree(L.h_out);
        (void)hipHostFree(L.h_desc);
    }
    for (auto &kv : S->pools) (void)hipFree(kv.second.d_slots);
    delete S;
}

}  // extern "C"
