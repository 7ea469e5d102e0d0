/* mcb_common.h - layouts shared by the host side (mcb_world.cpp, mcb_api.cpp) and the HIP kernels
 * (mcb_kernels.hip) of the MI355X biome engine.  Plain C structs, identical on host and device.
 *
 * Seed-independent "world" constants (one per version x dimension x preset) live in one device
 * buffer (McbWorld) read through the scalar cache; per-seed noise state lives in a slot array
 * (McbSlot) in HBM and is staged into LDS per workgroup.
 */
#ifndef MCB_COMMON_H
#define MCB_COMMON_H
#include <stdint.h>

#define MCB_MAX_LEV 12          /* levels per NormalNoise stack (continentalness has 9) */
#define MCB_OW_OCT 46           /* octaves of one Overworld seed: 23 levels x 2 stacks (every preset) */
#define MCB_NE_OCT 8            /* Nether: 2 noises x 2 stacks x 2 levels */
/* Levels per noise (offset, temperature, vegetation, continentalness, erosion, ridges): the same in every
 * supported version and preset (climate-26.x.txt); the kernels unroll them, the loader checks them. */
#define MCB_OW_LEVELS {3, 2, 2, 9, 4, 3}
#define MCB_NE_LEVELS {2, 2}
#define MCB_SPL_NODES 320
#define MCB_SPL_PTS 320
#define MCB_SPL_MAXD 6          /* device spline recursion levels: splines nest at most 5 deep */
#define MCB_CELL_MAX_BP 24      /* breakpoints per climate dimension in the cell table */
#define MCB_CELL_LUT (6 * 256)  /* cell locator entries */

enum { MCB_V26_1 = 0, MCB_V26_2 = 1, MCB_V26_3 = 2, MCB_V26_4 = 3 };   /* MCB_V26_4 = 26.4-snapshot-2 */
enum { MCB_OVERWORLD = 0, MCB_NETHER = 1, MCB_END = 2 };

/* One NormalNoise: per-level constants, engine/mc_noise.h:142-155 */
typedef struct {
    int32_t nlev, legacy, octs, first;
    int32_t lvl[MCB_MAX_LEV];
    double freq[MCB_MAX_LEV];       /* first-stack frequency of each level */
    double amp_d[MCB_MAX_LEV];      /* 26.1/26.2 amplitude */
    double vf_d[MCB_MAX_LEV];       /* 26.1/26.2 per-level value factor */
    double value_factor;            /* 26.1/26.2 NormalNoise value factor */
    float amp_f[MCB_MAX_LEV];       /* 26.3 layer amplitude */
    uint64_t hlo[MCB_MAX_LEV], hhi[MCB_MAX_LEV];   /* md5("octave_<n>") */
} McbNoise;

/* Offset spline (float CubicSpline), flattened. */
typedef struct {
    int32_t n_nodes, n_pts;
    int32_t kind[MCB_SPL_NODES];    /* 0 constant, 1 multipoint */
    int32_t coord[MCB_SPL_NODES];
    int32_t first[MCB_SPL_NODES];
    int32_t count[MCB_SPL_NODES];
    float value[MCB_SPL_NODES];
    float loc[MCB_SPL_PTS], der[MCB_SPL_PTS];
    int32_t child[MCB_SPL_PTS];
    uint32_t hdr[MCB_SPL_NODES];    /* packed for the device: multipoint << 31 | coord << 24 | count << 16 | first */
    int32_t depth;                  /* multipoint levels on the deepest path */
} McbSpline;

/* R-tree node in DFS pre-order: the first child of node i is i+1, and `escape` is the index just past
 * i's subtree, so the search walks the array without a stack.  Bounds are the quantized parameter
 * intervals of temperature..weirdness; the 7th (offset) dimension has target 0 and is folded into d6. */
typedef struct {
    int16_t lo[6], hi[6];
    uint32_t d6;                    /* squared distance of the offset interval to 0 */
    uint16_t escape;
    uint16_t leaf;                  /* 0x8000 | biome id for a leaf, 0 for a subtree */
} McbNode;

typedef struct {
    int32_t version, dim, fmt_new, n_noise;
    McbNoise noise[6];              /* overworld: offset,temp,veg,cont,eros,ridge; nether: temp, veg */
    uint64_t nlo[6], nhi[6];        /* md5("minecraft:<noise>") */
    int32_t oct_base[6];            /* first octave slot of each noise in a seed's state */
    int32_t n_oct;                  /* octaves per seed */
    int32_t gy0, gy1;
    double gv0_d, gv1_d, offc_d;
    float gv0_f, gv1_f, offc_f, grad_factor_f;
    double rf_d[3];
    float rf_f[3];
    int32_t end_ids[5];
    int32_t tree_n;                 /* nodes in the pre-order array (root = 0) */
    int32_t cell_on;                /* 1: Overworld cell table present (else every search walks the tree) */
    int32_t cell_stride[6];         /* cell index = sum of cell_stride[k] * (cell of dimension k) */
    /* cell locator, per tree dimension k and bucket b of 256 target values (t + 32768) >> 8:
     * (cells below the bucket's first value) * cell_stride[k] | (offset in the bucket of the one
     * breakpoint inside it, 256 = none) << 20 */
    uint32_t cell_lut[6][256];
    uint64_t jm262, ja262;          /* LCG jump by 262 draws (LegacyFbm skipped octave) */
    uint64_t jm_end, ja_end;        /* LCG jump by 17292 draws (End island noise) */
    McbSpline spline;
} McbWorld;

/* Per-octave record of a seed's state.  y- and z-derived constants serve the evaluations whose
 * y (climate noises, ShiftA) or z (ShiftB) coordinate is the octave offset itself. */
typedef struct {
    double xo, yo, zo;
    double fy_d, wy_d, fz_d, wz_d;  /* fraction and fade, double path */
    int32_t yi, zi;
    float fy_f, wy_f, fz_f, wz_f;   /* fraction and fade, float path */
} McbOct;                           /* 80 bytes */

/* Seed slot in HBM: permutation tables first (contiguous, staged to LDS), then octave records. */
#define MCB_SLOT_PERM_BYTES (MCB_OW_OCT * 256)
typedef struct {
    uint8_t perm[MCB_OW_OCT][256];
    McbOct oct[MCB_OW_OCT];
    int64_t seed;
    int64_t pad;
} McbSlot;                          /* 11776 + 3680 + 16 = 15472 bytes */

/* Work descriptors of the batched kernels. */
typedef struct {
    int32_t slot;                   /* seed slot */
    int32_t x0, z0;                 /* origin: quart (grid) or block (zoom) coordinates */
    int32_t nx, nz;                 /* extent */
    int32_t ny;                     /* layers */
    int32_t y[8];                   /* quart y (grid) or block y (zoom) of each layer */
    int32_t zoom;                   /* 1: block resolution through BiomeManager */
    int32_t world;                  /* index into the world table of a mixed launch */
    int64_t out_off;                /* output offset (bytes); layer j at out_off + j*nx*nz */
    int64_t hashed_seed;            /* BiomeManager seed (zoom only) */
} McbRequest;

typedef struct {
    int32_t req;                    /* request index */
    int16_t tx, tz;                 /* tile coordinates inside the request */
} McbTile;

#endif
