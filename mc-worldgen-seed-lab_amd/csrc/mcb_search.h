// mcb_search.h - seed-search launch arguments shared by the host (mcb_search.cpp) and the kernels
// (mcb_search.hip).
#ifndef MCB_SEARCH_H
#define MCB_SEARCH_H
#include "mcb_internal.h"

namespace mcb {

struct SearchArgs {
    const McbWorld *W;
    const McbNode *tree;
    const uint32_t *cells;          // Overworld cell table (entries, listed leaves)
    const McbNode *crecs;
    const uint4 *feas;              // biome sets of prefix cells (temperature, +vegetation, +continents, +erosion)
    int feas_off[4];
    int ncell[6];                   // cells per tree dimension
    const uint32_t *coord_lut;      // per dimension and bucket: cells below | breakpoint offset << 8
    const McbObs *obs;
    int nobs, flags;
    int src_kind;
    const int64_t *list;            // device copy of this chunk of a list source
    uint64_t base, n, first;        // candidates [first, first + n) of the source
    uint64_t *next;                 // work counter (candidates handed out in this chunk)
    int64_t *found;
    uint32_t *nfound;
    uint32_t max_found;
    unsigned long long *hist;       // nobs + 1 counters
};

hipError_t launch_search_ow(const DevWorld &w, const SearchArgs &A, int blocks, hipStream_t stream);
hipError_t launch_search_slots(const DevWorld &w, const SearchArgs &A, const McbSlot *d_slots, int ncand,
                               hipStream_t stream);
hipError_t launch_seed_gen(const SearchArgs &A, int n, int64_t *d_out, hipStream_t stream);

}  // namespace mcb
#endif
