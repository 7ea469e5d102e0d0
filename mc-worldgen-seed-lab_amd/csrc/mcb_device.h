// mcb_device.h - device functions shared by the HIP kernels of the biome engine (mcb_kernels.hip,
// mcb_search.hip): integer generators, octave evaluation, NormalNoise, offset spline, biome search
// (cell table and tree walk), climate columns, End islands, BiomeManager zoom steps.
//
// Precision follows the reference CPU engine exactly (engine/mc_noise.h, engine/mc_climate.h): 26.1/26.2
// noise in double, 26.3 noise with double coordinates and float arithmetic; built with -ffp-contract=off
// and no fast-math flags, so every multiply and add is a separate IEEE operation as in the reference
// (explicit fused multiply-adds appear only where the product is exact).
#ifndef MCB_DEVICE_H
#define MCB_DEVICE_H
#include "mcb_internal.h"

namespace mcb {
namespace dev {

// ================================================================================================
// integer generators (engine/mc_rng.h:11-131)
// ================================================================================================
struct Xo { uint64_t a, b; };

__device__ __forceinline__ uint64_t rotl64(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
__device__ __forceinline__ uint64_t mix13(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ Xo xo_make(uint64_t a, uint64_t b) {
    Xo r{a, b};
    if ((a | b) == 0) { r.a = 0x9E3779B97F4A7C15ULL; r.b = 0x6A09E667F3BCC909ULL; }
    return r;
}
__device__ __forceinline__ Xo xo_from_seed(int64_t seed) {
    uint64_t lo = (uint64_t)seed ^ 0x6A09E667F3BCC909ULL;
    return xo_make(mix13(lo), mix13(lo + 0x9E3779B97F4A7C15ULL));
}
__device__ __forceinline__ uint64_t xo_next(Xo &r) {
    uint64_t s0 = r.a, s1 = r.b;
    uint64_t out = rotl64(s0 + s1, 17) + s0;
    s1 ^= s0;
    r.a = rotl64(s0, 49) ^ s1 ^ (s1 << 21);
    r.b = rotl64(s1, 28);
    return out;
}
__device__ __forceinline__ int xo_bound(Xo &r, uint32_t bound) {
    uint64_t m = (uint64_t)(uint32_t)xo_next(r) * bound;
    if ((uint32_t)m < bound) {
        uint32_t thr = (0u - bound) % bound;
        while ((uint32_t)m < thr) m = (uint64_t)(uint32_t)xo_next(r) * bound;
    }
    return (int)(m >> 32);
}
__device__ __forceinline__ double xo_double(Xo &r) { return (double)(xo_next(r) >> 11) * 1.1102230246251565e-16; }
__device__ __forceinline__ Xo xo_fork(Xo &r) { Xo p; p.a = xo_next(r); p.b = xo_next(r); return p; }

struct Lcg { uint64_t s; };
#define LCG_MASK ((1ULL << 48) - 1)
__device__ __forceinline__ int32_t lcg_bits(Lcg &r, int bits) {
    r.s = (r.s * 0x5DEECE66DULL + 0xBULL) & LCG_MASK;
    return (int32_t)(int64_t)(r.s >> (48 - bits));
}
__device__ __forceinline__ int lcg_bound(Lcg &r, int32_t bound) {
    if ((bound & -bound) == bound) return (int32_t)(((int64_t)bound * (int64_t)lcg_bits(r, 31)) >> 31);
    for (;;) {
        int32_t bits = lcg_bits(r, 31), val = bits % bound;
        if ((int32_t)((uint32_t)bits - (uint32_t)val + (uint32_t)(bound - 1)) >= 0) return val;
    }
}
__device__ __forceinline__ double lcg_double(Lcg &r) {
    int64_t hi = lcg_bits(r, 26), lo = lcg_bits(r, 27);
    return (double)((hi << 27) + lo) * 1.1102230246251565e-16;
}

// ================================================================================================
// per-lane permutation tile in LDS: byte i of lane L sits in word (i>>2)*64 + L, so every lane's
// access hits its own bank column whatever i is (conflict-free Fisher-Yates).
// ================================================================================================
__device__ __forceinline__ uint8_t *tile_byte(uint32_t *tile, int lane, int i) {
    return reinterpret_cast<uint8_t *>(tile + (i >> 2) * 64 + lane) + (i & 3);
}

template <class Draw>
__device__ void fisher_yates(uint32_t *tile, int lane, Draw draw) {
    for (int w = 0; w < 64; w++) tile[w * 64 + lane] = (uint32_t)(4 * w) | (uint32_t)(4 * w + 1) << 8 | (uint32_t)(4 * w + 2) << 16 | (uint32_t)(4 * w + 3) << 24;
    for (int i = 0; i < 256; i++) {
        int j = i + draw(256 - i);
        uint8_t *pi = tile_byte(tile, lane, i), *pj = tile_byte(tile, lane, j);
        uint8_t a = *pi, b = *pj;
        *pi = b;
        *pj = a;
    }
}

__device__ __forceinline__ double fade_d(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }
__device__ __forceinline__ float fade_f(float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }

// Octave record: offsets plus the floor, fraction and fade of the y and z offsets (the coordinates of
// the climate noises' y and of ShiftB's z are the offsets themselves).
__device__ __forceinline__ McbOct make_oct(double xo, double yo, double zo) {
    McbOct r;
    r.xo = xo; r.yo = yo; r.zo = zo;
    double yf = floor(yo), zf = floor(zo);
    r.yi = (int)yf; r.zi = (int)zf;
    r.fy_d = yo - yf; r.fz_d = zo - zf;
    r.wy_d = fade_d(r.fy_d); r.wz_d = fade_d(r.fz_d);
    r.fy_f = (float)(yo - yf); r.fz_f = (float)(zo - zf);
    r.wy_f = fade_f(r.fy_f); r.wz_f = fade_f(r.fz_f);
    return r;
}

// Write one octave (permutation + derived constants) of a seed slot.
__device__ void store_octave(McbSlot *slot, int o, const uint32_t *tile, int lane, double xo, double yo, double zo) {
    uint4 *dst = reinterpret_cast<uint4 *>(slot->perm[o]);
    for (int q = 0; q < 16; q++)
        dst[q] = make_uint4(tile[(4 * q) * 64 + lane], tile[(4 * q + 1) * 64 + lane], tile[(4 * q + 2) * 64 + lane],
                            tile[(4 * q + 3) * 64 + lane]);
    slot->oct[o] = make_oct(xo, yo, zo);
}

// ================================================================================================
// octave evaluation (engine/mc_noise.h:54-110) on permutation tables in LDS
// ================================================================================================
// The grid and points kernels stage each octave's 256-entry permutation twice in a row (512 bytes),
// so P2[i] = perm[i & 255] for every index the lattice hash forms (a + y + 1 <= 511): no masks, and the
// +1 neighbours are immediate offsets of the same address.  The End's simplex noise keeps the plain
// 256-byte table.
#define PERM(i) ((int)P[(i) & 255])
#define OCT_BYTES 512

// Permutation accessors for the octave code: P2 = the doubled 512-byte table (indices up to 511, no
// mask); TileP = a lane's own table inside the Fisher-Yates tile (byte i in word (i>>2)*64 + lane).
struct P2 {
    const uint8_t *p;
    __device__ __forceinline__ int operator[](int i) const { return p[i]; }
};
struct TileP {
    const uint8_t *p;                                   // tile bytes + 4 * lane
    __device__ __forceinline__ int operator[](int i) const { return p[((i & 255) >> 2) * 256 + (i & 3)]; }
};
// GP = a 256-byte table in global memory (a seed slot's), masked.
struct GP {
    const uint8_t *p;
    __device__ __forceinline__ int operator[](int i) const { return p[i & 255]; }
};
// Permutation sources of the climate columns: the doubled LDS tables (a byte pointer, OCT_BYTES per
// octave) or a seed slot's own tables in global memory (SlotPerm).
struct SlotPerm {
    const uint8_t (*t)[256];
};
__device__ __forceinline__ P2 perm_at(const uint8_t *perm2, int o) { return P2{perm2 + o * OCT_BYTES}; }
__device__ __forceinline__ GP perm_at(SlotPerm s, int o) { return GP{s.t[o]}; }

// gradient table of ImprovedNoise (16 x {gx,gy,gz}), each component a 2-bit two's complement field
constexpr int kGrad[16][3] = {{1, 1, 0},  {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0}, {1, 0, 1},  {-1, 0, 1}, {1, 0, -1}, {-1, 0, -1},
                              {0, 1, 1},  {0, -1, 1}, {0, 1, -1}, {0, -1, -1}, {1, 1, 0},  {0, -1, 1}, {-1, 1, 0}, {0, -1, -1}};
constexpr uint32_t grad_bits(int c) {
    uint32_t v = 0;
    for (int h = 0; h < 16; h++) v |= (uint32_t)(kGrad[h][c] & 3) << (2 * h);
    return v;
}
constexpr uint32_t kGX = grad_bits(0), kGY = grad_bits(1), kGZ = grad_bits(2);
__device__ __forceinline__ int gcomp(uint32_t bits, int h) { return __builtin_amdgcn_sbfe((int)bits, 2 * h, 2); }

// LDS gradient entry for a permutation value h (the vector of h & 15), 8 bytes: gx as f32 in the low
// word, gy and gz as f16 (exact for -1, 0, 1) in the high word, so one ds_read_b64 feeds the dot product
// through mixed-precision FMAs.
typedef uint64_t GradEnt;
typedef _Float16 half2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void init_grad(GradEnt *G) {
    for (int h = threadIdx.x; h < 256; h += blockDim.x) {
        const int k = h & 15;
        const half2v yz = {(_Float16)gcomp(kGY, k), (_Float16)gcomp(kGZ, k)};
        G[h] = (uint64_t)__float_as_uint((float)gcomp(kGX, k)) | (uint64_t)__builtin_bit_cast(uint32_t, yz) << 32;
    }
}

// g . (x, y, z) as the reference's gx*x + gy*y + gz*z.  With g in {-1, 0, 1} every product is exact, so
// a fused multiply-add rounds exactly like the separate product and sum (same value, same signed zero).
__device__ __forceinline__ float gdot_f(GradEnt g, float x, float y, float z) {
    const float gx = __uint_as_float((uint32_t)g);
    const half2v yz = __builtin_bit_cast(half2v, (uint32_t)(g >> 32));
    return fmaf((float)yz.y, z, fmaf((float)yz.x, y, gx * x));
}
// Double-precision entry (26.1/26.2), 16 bytes read with one ds_read_b128: gx as a double, gy and gz as
// f32 (one conversion each).  Bit-field extraction plus integer conversion of all three components took
// eight VALU instructions per corner; this takes three (the index shift and two conversions).
struct __attribute__((aligned(16))) GradD {
    double gx;
    float gy, gz;
};
__device__ __forceinline__ void init_grad(GradD *G) {
    for (int h = threadIdx.x; h < 256; h += blockDim.x) {
        const int k = h & 15;
        G[h] = GradD{(double)gcomp(kGX, k), (float)gcomp(kGY, k), (float)gcomp(kGZ, k)};
    }
}
// gradient table type of a precision mode: 1 = float arithmetic (26.3), 0 = double (26.1/26.2)
template <int MODE> struct GT { typedef GradEnt T; };
template <> struct GT<0> { typedef GradD T; };
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ double gdot_d(const GradD &g, double x, double y, double z) {
    const u32x4 v = *reinterpret_cast<const u32x4 *>(&g);          // one 16-byte LDS read
    const double gx = __builtin_bit_cast(double, (uint64_t)v.y << 32 | v.x);
    return fma((double)__uint_as_float(v.w), z, fma((double)__uint_as_float(v.z), y, gx * x));
}
__device__ __forceinline__ float lerp_f(float t, float a, float b) { return a + t * (b - a); }
__device__ __forceinline__ double lerp_d(double t, double a, double b) { return a + t * (b - a); }

// PerlinNoise.wrap: identity inside +-2^24, else x - floor(x/2^25 + 0.5)*2^25 (mc_noise.h:77,86-89)
// WRAP = false: the caller has shown that no input reaches +-2^24 (wrap is then the identity), and the
// per-coordinate test and its branch are left out.
template <bool WRAP = true>
__device__ __forceinline__ double wrapc(double x) {
    if (WRAP && __builtin_expect(!(fabs(x) < 16777216.0), 0)) x = x - floor(x * (1.0 / 33554432.0) + 0.5) * 33554432.0;
    return x;
}

template <int MODE> struct R;
template <> struct R<1> { typedef float T; };
template <> struct R<0> { typedef double T; };

// Trilinear blend of the 8 lattice corners (the reference's nesting: x, then y, then z).  hXYZ = hash
// of corner (X, Y, Z); f* = fractions; u, v, w = fades along x, y, z.
template <int MODE>
__device__ __forceinline__ typename R<MODE>::T corners(const typename GT<MODE>::T *G, int h000, int h100, int h010, int h110, int h001,
                                                       int h101, int h011, int h111, typename R<MODE>::T fx,
                                                       typename R<MODE>::T fy, typename R<MODE>::T fz,
                                                       typename R<MODE>::T u, typename R<MODE>::T v, typename R<MODE>::T w) {
    if constexpr (MODE == 1) {
        const float gx = fx - 1.0f, gy = fy - 1.0f, gz = fz - 1.0f;
        const float c000 = gdot_f(G[h000], fx, fy, fz), c100 = gdot_f(G[h100], gx, fy, fz);
        const float c010 = gdot_f(G[h010], fx, gy, fz), c110 = gdot_f(G[h110], gx, gy, fz);
        const float c001 = gdot_f(G[h001], fx, fy, gz), c101 = gdot_f(G[h101], gx, fy, gz);
        const float c011 = gdot_f(G[h011], fx, gy, gz), c111 = gdot_f(G[h111], gx, gy, gz);
        return lerp_f(w, lerp_f(v, lerp_f(u, c000, c100), lerp_f(u, c010, c110)),
                      lerp_f(v, lerp_f(u, c001, c101), lerp_f(u, c011, c111)));
    } else {
        const double gx = fx - 1.0, gy = fy - 1.0, gz = fz - 1.0;
        const double c000 = gdot_d(G[h000], fx, fy, fz), c100 = gdot_d(G[h100], gx, fy, fz);
        const double c010 = gdot_d(G[h010], fx, gy, fz), c110 = gdot_d(G[h110], gx, gy, fz);
        const double c001 = gdot_d(G[h001], fx, fy, gz), c101 = gdot_d(G[h101], gx, fy, gz);
        const double c011 = gdot_d(G[h011], fx, gy, gz), c111 = gdot_d(G[h111], gx, gy, gz);
        return lerp_d(w, lerp_d(v, lerp_d(u, c000, c100), lerp_d(u, c010, c110)),
                      lerp_d(v, lerp_d(u, c001, c101), lerp_d(u, c011, c111)));
    }
}

// Octave whose y coordinate is the octave offset (climate noises, ShiftA): x, z vary.
template <int MODE, class PA, bool WRAP = true>
__device__ __forceinline__ typename R<MODE>::T oct_yc(PA P2, const typename GT<MODE>::T *G, const McbOct &o, double X, double Z) {
    typedef typename R<MODE>::T T;
    const double x = wrapc<WRAP>(X) + o.xo, z = wrapc<WRAP>(Z) + o.zo;
    const double xf = floor(x), zf = floor(z);
    const int xi = (int)xf & 255, zi = (int)zf & 255, yi = o.yi & 255;
    const T fx = (T)(x - xf), fz = (T)(z - zf);
    const T fy = MODE ? (T)o.fy_f : (T)o.fy_d, wy = MODE ? (T)o.wy_f : (T)o.wy_d;
    const int a = P2[xi], b = P2[xi + 1];
    const int aa = P2[a + yi], ab = P2[a + yi + 1], ba = P2[b + yi], bb = P2[b + yi + 1];
    const T u = MODE ? (T)fade_f((float)fx) : (T)fade_d((double)fx);
    const T w = MODE ? (T)fade_f((float)fz) : (T)fade_d((double)fz);
    return corners<MODE>(G, P2[aa + zi], P2[ba + zi], P2[ab + zi], P2[bb + zi], P2[aa + zi + 1], P2[ba + zi + 1],
                         P2[ab + zi + 1], P2[bb + zi + 1], fx, fy, fz, u, wy, w);
}

// Octave whose z coordinate is the octave offset (ShiftB): x, y vary.
template <int MODE, class PA, bool WRAP = true>
__device__ __forceinline__ typename R<MODE>::T oct_zc(PA P2, const typename GT<MODE>::T *G, const McbOct &o, double X, double Y) {
    typedef typename R<MODE>::T T;
    const double x = wrapc<WRAP>(X) + o.xo, y = wrapc<WRAP>(Y) + o.yo;
    const double xf = floor(x), yf = floor(y);
    const int xi = (int)xf & 255, yi = (int)yf & 255, zi = o.zi & 255;
    const T fx = (T)(x - xf), fy = (T)(y - yf);
    const T fz = MODE ? (T)o.fz_f : (T)o.fz_d, wz = MODE ? (T)o.wz_f : (T)o.wz_d;
    const int a = P2[xi], b = P2[xi + 1];
    const int aa = P2[a + yi], ab = P2[a + yi + 1], ba = P2[b + yi], bb = P2[b + yi + 1];
    const T u = MODE ? (T)fade_f((float)fx) : (T)fade_d((double)fx);
    const T v = MODE ? (T)fade_f((float)fy) : (T)fade_d((double)fy);
    return corners<MODE>(G, P2[aa + zi], P2[ba + zi], P2[ab + zi], P2[bb + zi], P2[aa + zi + 1], P2[ba + zi + 1],
                         P2[ab + zi + 1], P2[bb + zi + 1], fx, fy, fz, u, v, wz);
}

#define INPUT_FACTOR 1.0181268882175227

// One level of one stack of a NormalNoise at (X, Y) (y or z fixed at the octave offset): its term of the
// reference's sum.  26.3: amp_f[k] * octave(X*f, Y*f) with f = freq[k], or freq[k]*1.0181... for the
// second stack (NoiseStack.get, mc_noise.h:303-321); 26.1/26.2: amp_d[k] * octave(wrap(X'*f), wrap(Y'*f))
// * vf_d[k] with X' = X, or X*1.0181... for the second stack (mc_noise.h:286-301).
template <int MODE, bool ZC, class PA, bool WRAP = true>
__device__ __forceinline__ typename R<MODE>::T octave_term(PA P, const typename GT<MODE>::T *G, const McbOct &o, const McbNoise &N, int k,
                                                           int st, double X, double Y) {
    if constexpr (MODE == 1) {
        const double f = st ? N.freq[k] * INPUT_FACTOR : N.freq[k];
        const float v = ZC ? oct_zc<1, PA, WRAP>(P, G, o, X * f, Y * f) : oct_yc<1, PA, WRAP>(P, G, o, X * f, Y * f);
        return N.amp_f[k] * v;
    } else {
        const double Xs = st ? X * INPUT_FACTOR : X, Ys = st ? Y * INPUT_FACTOR : Y, f = N.freq[k];
        // the octave's own wrap is the identity on an already wrapped input (|x| <= 2^24): left out
        const double v = ZC ? oct_zc<0, PA, false>(P, G, o, wrapc<WRAP>(Xs * f), wrapc<WRAP>(Ys * f))
                            : oct_yc<0, PA, false>(P, G, o, wrapc<WRAP>(Xs * f), wrapc<WRAP>(Ys * f));
        return N.amp_d[k] * v * N.vf_d[k];
    }
}

// NormalNoise with y fixed at the octave offsets, NL levels whose octaves start at slot BASE (the
// layout the loader checks), fully unrolled so every table and octave record sits at a constant
// offset.  26.3 interleaves first/second layers and accumulates in float (NoiseStack.get,
// mc_noise.h:303-312; legacy-nether order :313-321); 26.1/26.2 sums each stack in double, then
// (v1+v2)*valueFactor (mc_noise.h:286-301).
template <int MODE, bool ZC, int NL, int BASE, bool LEGACY, class PS, bool WRAP = true>
__device__ __forceinline__ typename R<MODE>::T normal2(PS perm2, const typename GT<MODE>::T *G, const McbOct *oct,
                                                       const McbNoise &N, double X, double Y) {
    auto term = [&](int k, int st) {
        const int o = BASE + st * NL + k;
        return octave_term<MODE, ZC, decltype(perm_at(perm2, o)), WRAP>(perm_at(perm2, o), G, oct[o], N, k, st, X, Y);
    };
    if constexpr (MODE == 1) {
        float v = 0.0f;
        if constexpr (!LEGACY) {
#pragma unroll 1
            for (int k = 0; k < NL; k++) {
                v += term(k, 0);
                v += term(k, 1);
            }
        } else {
#pragma unroll 1
            for (int k = 0; k < NL; k++) v += term(k, 0);
#pragma unroll 1
            for (int k = 0; k < NL; k++) v += term(k, 1);
        }
        return v;
    } else {
        double v1 = 0.0, v2 = 0.0;
#pragma unroll 1
        for (int k = 0; k < NL; k++) v1 += term(k, 0);
#pragma unroll 1
        for (int k = 0; k < NL; k++) v2 += term(k, 1);
        return (v1 + v2) * N.value_factor;
    }
}

// ================================================================================================
// offset spline (engine/mc_climate.h:40-66), float
// ================================================================================================
// The spline lives in LDS (packed headers, locations, derivatives, children).  Evaluation recurses
// through MCB_SPL_MAXD compile-time levels; each level has one call site (a loop over the one or two
// children the node needs), so the code grows linearly with the depth.
struct SplLds {
    uint32_t hdr[MCB_SPL_NODES];
    float val[MCB_SPL_NODES];
    float loc[MCB_SPL_PTS];
    float der[MCB_SPL_PTS];
    uint16_t child[MCB_SPL_PTS];
};
__device__ __forceinline__ void stage_spline(SplLds *L, const McbSpline &S) {
    for (int i = threadIdx.x; i < MCB_SPL_NODES; i += blockDim.x) {
        L->hdr[i] = S.hdr[i];
        L->val[i] = S.value[i];
    }
    for (int i = threadIdx.x; i < MCB_SPL_PTS; i += blockDim.x) {
        L->loc[i] = S.loc[i];
        L->der[i] = S.der[i];
        L->child[i] = (uint16_t)S.child[i];
    }
}
template <int D>
__device__ __forceinline__ float spline_eval(const SplLds &S, int id, float c, float e, float w, float r) {
    if constexpr (D == 0) {
        return __builtin_nanf("");                     // unreachable: the host bounds the nesting depth
    } else {
        const uint32_t h = S.hdr[id];
        if (!(h >> 31)) return S.val[id];
        const int first = (int)(h & 0xFFFFu), cnt = (int)((h >> 16) & 0xFFu), coord = (int)((h >> 24) & 3u);
        const float x = coord == 0 ? c : (coord == 1 ? e : (coord == 2 ? w : r));
        int lo = 0, hi = cnt;                          // first i with x < loc[i], minus one
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (x < S.loc[first + mid]) hi = mid; else lo = mid + 1;
        }
        const int i = lo - 1, last = cnt - 1;
        const bool edge = i < 0 || i == last;          // linear extension beyond the end points
        const int base = first + (i < 0 ? 0 : i);
        float ya = 0.0f, yb = 0.0f;
#pragma unroll 1
        for (int q = 0; q < (edge ? 1 : 2); q++) {
            const float v = spline_eval<D - 1>(S, S.child[base + q], c, e, w, r);
            if (q == 0) ya = v; else yb = v;
        }
        if (edge) {
            const float d = S.der[base];
            return d == 0.0f ? ya : ya + d * (x - S.loc[base]);
        }
        const float x1 = S.loc[base], x2 = S.loc[base + 1];
        const float t = (x - x1) / (x2 - x1);
        const float a = S.der[base] * (x2 - x1) - (yb - ya);
        const float b = -S.der[base + 1] * (x2 - x1) + (yb - ya);
        return lerp_f(t, ya, yb) + t * (1.0f - t) * lerp_f(t, a, b);
    }
}

// ================================================================================================
// R-tree search (engine/mc_biomes.h:31-62), exact: the result is the leaf minimising
// (fitness, pre-order index) -- the first minimum in the reference's child order.  The array is in
// pre-order with escape links, so the walk needs no stack; a warm-start candidate (a nearby sample's
// leaf) only tightens the initial bound and never changes the result.
// ================================================================================================
// squared distance of one parameter interval to the target: (t - clamp(t, lo, hi))^2, integer-valued
// doubles (exact below 2^53)
// One dimension's term of the fitness: the game's above = t - hi, below = lo - t, squared
// (engine/mc_biomes.h:31-39).  max(above, below, 0) equals it for every box with lo <= hi + 1, which
// covers 26.4-snapshot-2's boxes (lo = hi + 1 after its exclusive upper bound) as well as closed ones.
__device__ __forceinline__ double dim_sq(int t, int lo, int hi, double acc) {
    const double x = (double)max(max(t - hi, lo - t), 0);
    return fma(x, x, acc);
}
__device__ __forceinline__ double node_dist(const McbNode &nd, const int *t) {
    double d = (double)nd.d6;
#pragma unroll
    for (int k = 0; k < 6; k++) d = dim_sq(t[k], nd.lo[k], nd.hi[k], d);
    return d;
}
__device__ __forceinline__ int tree_search(const McbNode *__restrict__ N, int n, const int *t, int &cand) {
    int bi = n;
    double bd = __builtin_inf();
    if (cand > 0) { bi = cand; bd = node_dist(N[cand], t); }
    int cur = 1;
    while (cur < n) {
        const McbNode nd = N[cur];
        double cd = node_dist(nd, t);
        bool better = cd < bd || (cd == bd && cur < bi);
        bool leaf = nd.leaf != 0;
        if (better && leaf) { bd = cd; bi = cur; }
        cur = (better && !leaf) ? cur + 1 : (int)nd.escape;
    }
    cand = bi;
    return N[bi].leaf & 0xFF;
}

// All y layers of one column in one walk.  The layers share temperature..weirdness and the offset
// dimension, so each node's 5-dimension part is computed once and only the depth term per layer; a
// subtree is entered when any layer may improve in it.  Extra visits never change a layer's result
// (its own (distance, index) rule decides every update), so each layer gets exactly its
// single-search leaf.  cand[j] carries the previous column's leaf of layer j as warm start.
template <int MAXL>
__device__ __forceinline__ void tree_search_layers(const McbNode *__restrict__ N, int n, const int *t5, const int *dep,
                                                   int ny, int *cand, int *biome) {
    double bd[MAXL];
    int bi[MAXL];
#pragma unroll
    for (int j = 0; j < MAXL; j++) {
        bi[j] = n;
        bd[j] = __builtin_inf();
        if (j < ny && cand[j] > 0) {
            const McbNode c = N[cand[j]];
            double d = (double)c.d6;
            d = dim_sq(t5[0], c.lo[0], c.hi[0], d);
            d = dim_sq(t5[1], c.lo[1], c.hi[1], d);
            d = dim_sq(t5[2], c.lo[2], c.hi[2], d);
            d = dim_sq(t5[3], c.lo[3], c.hi[3], d);
            d = dim_sq(t5[4], c.lo[5], c.hi[5], d);
            bd[j] = dim_sq(dep[j], c.lo[4], c.hi[4], d);
            bi[j] = cand[j];
        }
    }
    int cur = 1;
    while (cur < n) {
        const McbNode nd = N[cur];
        double d5 = (double)nd.d6;
        d5 = dim_sq(t5[0], nd.lo[0], nd.hi[0], d5);
        d5 = dim_sq(t5[1], nd.lo[1], nd.hi[1], d5);
        d5 = dim_sq(t5[2], nd.lo[2], nd.hi[2], d5);
        d5 = dim_sq(t5[3], nd.lo[3], nd.hi[3], d5);
        d5 = dim_sq(t5[4], nd.lo[5], nd.hi[5], d5);
        const bool leaf = nd.leaf != 0;
        bool any = false;
#pragma unroll
        for (int j = 0; j < MAXL; j++) {
            if (j < ny) {
                const double cd = dim_sq(dep[j], nd.lo[4], nd.hi[4], d5);
                const bool better = cd < bd[j] || (cd == bd[j] && cur < bi[j]);
                if (better && leaf) { bd[j] = cd; bi[j] = cur; }
                any |= better;
            }
        }
        cur = (any && !leaf) ? cur + 1 : (int)nd.escape;
    }
#pragma unroll
    for (int j = 0; j < MAXL; j++)
        if (j < ny) { cand[j] = bi[j]; biome[j] = N[bi[j]].leaf & 0xFF; }
}

// ================================================================================================
// climate cell table (built by mcb_cells.cpp): the cell's ascending list of the leaves that can be the
// (distance, pre-order index) minimum inside it.  Evaluating the list with a strict < keeps the first
// minimum, which is the tree search's result.  Targets beyond +-2^17 walk the tree instead.
// ================================================================================================
#define CELL_TMAX (1 << 17)
__device__ __forceinline__ bool cell_in_range(int t) { return t >= -CELL_TMAX && t <= CELL_TMAX; }
// Offset of target t's cell along tree dimension k (see McbWorld.cell_lut): the bucket entry gives the
// cells below the bucket, plus one stride when t is at or past the bucket's breakpoint.  Every
// breakpoint lies in (-32768, 32767], so a target beyond that range belongs to an outer cell and is
// clamped onto it first.
__device__ __forceinline__ uint32_t cell_part(const uint32_t *lut, int k, int t, uint32_t stride) {
    const uint32_t u = (uint32_t)(min(max(t, -32768), 32767) + 32768);
    const uint32_t e = lut[k * 256 + (int)(u >> 8)];
    return (e & 0xFFFFFu) + ((u & 255u) >= (e >> 20) ? stride : 0u);
}
// t5 = temperature, humidity, continentalness, erosion, weirdness; depth separately (tree dimension 4)
__device__ __forceinline__ uint32_t cell_base(const McbWorld *__restrict__ W, const uint32_t *lut, const int *t5) {
    return cell_part(lut, 0, t5[0], W->cell_stride[0]) + cell_part(lut, 1, t5[1], W->cell_stride[1]) +
           cell_part(lut, 2, t5[2], W->cell_stride[2]) + cell_part(lut, 3, t5[3], W->cell_stride[3]) +
           cell_part(lut, 5, t5[4], W->cell_stride[5]);
}
__device__ __forceinline__ uint32_t cell_depth(const uint32_t *lut, int dep) { return cell_part(lut, 4, dep, 1u); }
__device__ __forceinline__ void stage_lut(uint32_t *lds, const McbWorld *__restrict__ W) {
    for (int i = threadIdx.x; i < MCB_CELL_LUT; i += blockDim.x) lds[i] = W->cell_lut[0][i];
}
// Device cell entry: count in bits 0-7 (0 = walk the tree); count 1 = one biome answers the whole cell
// (one leaf, or leaves of one biome): biome in bits 8-15; else the first of `count` leaf records
// (copies of the listed leaves, ascending pre-order index) in bits 8-31.
__device__ __forceinline__ double leaf_dist5(const McbNode &nd, const int *t5, int dep) {
    double d = (double)nd.d6;
    d = dim_sq(t5[0], nd.lo[0], nd.hi[0], d);
    d = dim_sq(t5[1], nd.lo[1], nd.hi[1], d);
    d = dim_sq(t5[2], nd.lo[2], nd.hi[2], d);
    d = dim_sq(t5[3], nd.lo[3], nd.hi[3], d);
    d = dim_sq(t5[4], nd.lo[5], nd.hi[5], d);
    return dim_sq(dep, nd.lo[4], nd.hi[4], d);
}
// biome of the least (distance, index) leaf of an entry (count > 0); strict < keeps the first minimum
__device__ __forceinline__ int cell_best(const McbNode *__restrict__ recs, uint32_t e, const int *t5, int dep) {
    const int cnt = (int)(e & 255u);
    if (cnt == 1) return (int)((e >> 8) & 255u);
    const McbNode *r = recs + (e >> 8);
    double bd = __builtin_inf();
    int biome = 0;
    for (int q = 0; q < cnt; q++) {
        const McbNode nd = r[q];
        const double d = leaf_dist5(nd, t5, dep);
        if (d < bd) {
            bd = d;
            biome = nd.leaf & 0xFF;
        }
    }
    return biome;
}
// All layers of a column at once: the entries are loaded together; a layer whose entry names one biome is
// settled by it.  The others scan the union list of the column's five-dimension cell (every leaf any of
// its depth cells lists, in ascending pre-order): each record is read once, whole (two 16-byte loads), its
// five-dimension distance summed once, and only the depth term added per layer.  A leaf of the union that
// a layer's own list leaves out can never be that layer's (distance, index) minimum, so the strict < over
// the ascending union keeps the list's answer.  Without a union (count 0) each layer scans its own list.
// False when some layer must walk the tree.
__device__ __forceinline__ void load_node(const McbNode *__restrict__ recs, uint32_t i, McbNode &nd) {
    const uint4 *r = reinterpret_cast<const uint4 *>(recs + i);
    const uint4 lo = r[0], hi = r[1];
    __builtin_memcpy(reinterpret_cast<uint4 *>(&nd), &lo, 16);
    __builtin_memcpy(reinterpret_cast<uint4 *>(&nd) + 1, &hi, 16);
}
template <int MAXL>
__device__ __forceinline__ bool cell_layers(const McbWorld *__restrict__ W, const uint32_t *clut,
                                            const uint32_t *__restrict__ cells, const McbNode *__restrict__ recs,
                                            const uint32_t *__restrict__ ucells, const McbNode *__restrict__ urecs,
                                            const int *t5, const int *dep, int ny, int *b) {
    const uint32_t c5 = cell_base(W, clut, t5);
    uint32_t e[MAXL];
    int cnt[MAXL], mx = 0;
    bool ok = true, open = false;
    // the union entry is read with the layers' entries, so both are in flight together
    const uint32_t u = ucells[c5 / (uint32_t)W->cell_stride[5]];
#pragma unroll
    for (int j = 0; j < MAXL; j++) e[j] = j < ny ? cells[c5 + cell_depth(clut, dep[j])] : 1u;
#pragma unroll
    for (int j = 0; j < MAXL; j++) {
        cnt[j] = (int)(e[j] & 255u);
        ok = ok && cnt[j] != 0;
        if (cnt[j] == 1) b[j] = (int)((e[j] >> 8) & 255u);
        open = open || cnt[j] > 1;
        mx = max(mx, cnt[j] == 1 ? 0 : cnt[j]);
    }
    if (!ok) return false;
    if (!open) return true;
    double bd[MAXL];
#pragma unroll
    for (int j = 0; j < MAXL; j++) bd[j] = __builtin_inf();
    const int un = (int)(u & 255u);
    if (un) {
        // record q + 1 is requested before record q is compared, so its load overlaps the arithmetic
        McbNode next;
        load_node(urecs, u >> 8, next);
        for (int q = 0; q < un; q++) {
            const McbNode nd = next;
            if (q + 1 < un) load_node(urecs, (u >> 8) + q + 1, next);
            double d5 = (double)nd.d6;
            d5 = dim_sq(t5[0], nd.lo[0], nd.hi[0], d5);
            d5 = dim_sq(t5[1], nd.lo[1], nd.hi[1], d5);
            d5 = dim_sq(t5[2], nd.lo[2], nd.hi[2], d5);
            d5 = dim_sq(t5[3], nd.lo[3], nd.hi[3], d5);
            d5 = dim_sq(t5[4], nd.lo[5], nd.hi[5], d5);
#pragma unroll
            for (int j = 0; j < MAXL; j++) {
                if (cnt[j] > 1) {
                    const double d = dim_sq(dep[j], nd.lo[4], nd.hi[4], d5);
                    if (d < bd[j]) {
                        bd[j] = d;
                        b[j] = nd.leaf & 0xFF;
                    }
                }
            }
        }
        return true;
    }
    // Step q reads record q of every layer at once (a layer with no record q re-reads record 0 and ignores
    // it), and only then compares, so the loads of all layers are in flight together.  Record 0 exists
    // whenever some entry lists two or more leaves (mx > 0).
    for (int q = 0; q < mx; q++) {
        McbNode nd[MAXL];
#pragma unroll
        for (int j = 0; j < MAXL; j++) load_node(recs, q < cnt[j] && cnt[j] > 1 ? (e[j] >> 8) + q : 0u, nd[j]);
#pragma unroll
        for (int j = 0; j < MAXL; j++) {
            if (q < cnt[j] && cnt[j] > 1) {
                const double d = leaf_dist5(nd[j], t5, dep[j]);
                if (d < bd[j]) {
                    bd[j] = d;
                    b[j] = nd[j].leaf & 0xFF;
                }
            }
        }
    }
    return true;
}

__device__ __forceinline__ int quantize(float v) { return (int)(v * 10000.0f); }   // mc_climate.h:93

// ================================================================================================
// climate of one Overworld column (engine/mc_climate.h:113-164); depth is finished per layer
// ================================================================================================
struct Column {
    int t, h, c, e, w;      // quantized temperature, humidity, continentalness, erosion, weirdness
    float off_f;            // 26.3 offset (float)
    double off_d;           // 26.1/26.2 offset (double)
};

// Quantized targets and offset of a column from its five noise values (engine/mc_climate.h:113-164):
// ridges folded, offset = constant + spline(continents, erosion, ridges, folded ridges).
template <int MODE>
__device__ __forceinline__ void column_finish(const McbWorld *__restrict__ W, const SplLds *SP, typename R<MODE>::T T,
                                              typename R<MODE>::T H, typename R<MODE>::T C, typename R<MODE>::T E,
                                              typename R<MODE>::T Wr, Column &c) {
    if constexpr (MODE == 1) {
        float a = fabsf(Wr) + W->rf_f[0];
        float rf = (fabsf(a) + W->rf_f[1]) * W->rf_f[2];
        c.off_f = W->offc_f + spline_eval<MCB_SPL_MAXD>(*SP, 0, C, E, Wr, rf);
        c.t = quantize(T); c.h = quantize(H); c.c = quantize(C); c.e = quantize(E); c.w = quantize(Wr);
    } else {
        double rf = (fabs(fabs(Wr) + W->rf_d[0]) + W->rf_d[1]) * W->rf_d[2];
        float sp = spline_eval<MCB_SPL_MAXD>(*SP, 0, (float)C, (float)E, (float)Wr, (float)rf);
        c.off_d = 0.0 + (W->offc_d + (double)sp);
        c.t = quantize((float)T); c.h = quantize((float)H); c.c = quantize((float)C); c.e = quantize((float)E);
        c.w = quantize((float)Wr);
    }
}
// The shifts of the climate coordinates: X + ShiftA*4, Z + ShiftB*4 with ShiftA/ShiftB the offset noise
// at (X, 0, Z) and (Z, X, 0)
template <int MODE>
__device__ __forceinline__ void shift_coords(typename R<MODE>::T a, typename R<MODE>::T b, double X, double Z, double &x,
                                             double &z) {
    if constexpr (MODE == 1) {
        const float sx = a * 4.0f, sz = b * 4.0f;
        x = X + (double)sx;
        z = Z + (double)sz;
    } else {
        const double sx = a * 4.0, sz = b * 4.0;
        x = X + sx;
        z = Z + sz;
    }
}

// The five climate noises of a column: temperature, humidity, continentalness, erosion, weirdness at the
// shifted coordinates (engine/mc_climate.h:113-164, before the spline and quantization)
template <int MODE> struct Noises5 { typename R<MODE>::T t, h, c, e, w; };
template <int MODE, class PS, bool WRAP = true>
__device__ __forceinline__ Noises5<MODE> ow_noises(const McbWorld *__restrict__ W, PS perm, const typename GT<MODE>::T *G,
                                                   const McbOct *oct, int qx, int qz) {
    typedef typename R<MODE>::T T_;
    const McbNoise *N = W->noise;
    const double X = (double)qx, Z = (double)qz;     // blockX*0.25 with blockX = 4*quartX
    // octave slots of the layout MCB_OW_LEVELS {3, 2, 2, 9, 4, 3}: bases 0, 6, 10, 14, 32, 40
    const T_ a = normal2<MODE, false, 3, 0, false, PS, WRAP>(perm, G, oct, N[0], X, Z);
    const T_ b = normal2<MODE, true, 3, 0, false, PS, WRAP>(perm, G, oct, N[0], Z, X);
    double x, z;
    shift_coords<MODE>(a, b, X, Z, x, z);
    Noises5<MODE> r;
    r.t = normal2<MODE, false, 2, 6, false, PS, WRAP>(perm, G, oct, N[1], x, z);
    r.h = normal2<MODE, false, 2, 10, false, PS, WRAP>(perm, G, oct, N[2], x, z);
    r.c = normal2<MODE, false, 9, 14, false, PS, WRAP>(perm, G, oct, N[3], x, z);
    r.e = normal2<MODE, false, 4, 32, false, PS, WRAP>(perm, G, oct, N[4], x, z);
    r.w = normal2<MODE, false, 3, 40, false, PS, WRAP>(perm, G, oct, N[5], x, z);
    return r;
}
template <int MODE, class PS, bool WRAP = true>
__device__ __forceinline__ void ow_column(const McbWorld *__restrict__ W, PS perm, const typename GT<MODE>::T *G,
                                          const SplLds *SP, const McbOct *oct, int qx, int qz, Column &c) {
    const Noises5<MODE> n = ow_noises<MODE, PS, WRAP>(W, perm, G, oct, qx, qz);
    column_finish<MODE>(W, SP, n.t, n.h, n.c, n.e, n.w, c);
}

// depth = clamped y gradient at block y + offset, as the float the sampler returns
template <int MODE>
__device__ __forceinline__ float ow_depth_raw(const McbWorld *__restrict__ W, const Column &c, int by) {
    if (MODE) {
        int lo = min(W->gy0, W->gy1), hi = max(W->gy0, W->gy1);
        int yy = by < lo ? lo : (by > hi ? hi : by);
        float g = W->gv0_f + (float)(yy - W->gy0) * W->grad_factor_f;
        return g + c.off_f;
    } else {
        double fac = ((double)by - (double)W->gy0) / ((double)W->gy1 - (double)W->gy0);
        double g = fac < 0.0 ? W->gv0_d : (fac > 1.0 ? W->gv1_d : W->gv0_d + fac * (W->gv1_d - W->gv0_d));
        return (float)(g + c.off_d);
    }
}
template <int MODE>
__device__ __forceinline__ int ow_depth(const McbWorld *__restrict__ W, const Column &c, int qy) {
    return quantize(ow_depth_raw<MODE>(W, c, qy * 4));
}

// Nether climate: two legacy NormalNoises at (bx*0.25, by*0.0, bz*0.25) (engine/cuda/mcg_host.c:63-74)
template <int MODE, class PS, bool WRAP = true>
__device__ __forceinline__ void ne_column(const McbWorld *__restrict__ W, PS perm, const typename GT<MODE>::T *G,
                                          const McbOct *oct, int qx, int qz, int *t) {
    const double X = (double)qx, Z = (double)qz;
    // layout MCB_NE_LEVELS {2, 2}, legacy order: bases 0 and 4
    float T = (float)normal2<MODE, false, 2, 0, true, PS, WRAP>(perm, G, oct, W->noise[0], X, Z);
    float H = (float)normal2<MODE, false, 2, 4, true, PS, WRAP>(perm, G, oct, W->noise[1], X, Z);
    t[0] = quantize(T); t[1] = quantize(H);
    t[2] = t[3] = t[4] = t[5] = 0;
}

// ================================================================================================
// End islands (engine/mc_end.h:38-119)
// ================================================================================================
__device__ double simplex2(const uint8_t *P, double xin, double yin) {
    const double SQ3 = 1.7320508075688772;
    const double F2 = 0.5 * (SQ3 - 1.0), G2 = (3.0 - SQ3) / 6.0;
    double s = (xin + yin) * F2;
    int i = (int)floor(xin + s), j = (int)floor(yin + s);
    double t = (double)(i + j) * G2;
    double x0 = xin - ((double)i - t), y0 = yin - ((double)j - t);
    int i1 = x0 > y0 ? 1 : 0, j1 = 1 - i1;
    double x1 = x0 - (double)i1 + G2, y1 = y0 - (double)j1 + G2;
    double x2 = x0 - 1.0 + 2.0 * G2, y2 = y0 - 1.0 + 2.0 * G2;
    int ii = i & 255, jj = j & 255;
    int g0 = PERM(ii + PERM(jj)) % 12, g1 = PERM(ii + i1 + PERM(jj + j1)) % 12, g2 = PERM(ii + 1 + PERM(jj + 1)) % 12;
    double n0 = 0.0, n1 = 0.0, n2 = 0.0;
    double t0 = 0.5 - x0 * x0 - y0 * y0 - 0.0 * 0.0;
    if (!(t0 < 0.0)) { t0 *= t0; n0 = t0 * t0 * ((double)gcomp(kGX, g0) * x0 + (double)gcomp(kGY, g0) * y0 + (double)gcomp(kGZ, g0) * 0.0); }
    double t1 = 0.5 - x1 * x1 - y1 * y1 - 0.0 * 0.0;
    if (!(t1 < 0.0)) { t1 *= t1; n1 = t1 * t1 * ((double)gcomp(kGX, g1) * x1 + (double)gcomp(kGY, g1) * y1 + (double)gcomp(kGZ, g1) * 0.0); }
    double t2 = 0.5 - x2 * x2 - y2 * y2 - 0.0 * 0.0;
    if (!(t2 < 0.0)) { t2 *= t2; n2 = t2 * t2 * ((double)gcomp(kGX, g2) * x2 + (double)gcomp(kGY, g2) * y2 + (double)gcomp(kGZ, g2) * 0.0); }
    return 70.0 * (n0 + n1 + n2);
}
__device__ __forceinline__ float clampf_ref(float v, float lo, float hi) { float m = v < hi ? v : hi; return v < lo ? lo : m; }

// getHeightValue over the cell window +-rad around the section's chunk (rad = 12 in the game; the
// seed search may narrow it where the observed End biome set cannot depend on farther cells,
// engine/cuda/mc_gpu.cuh:606-650)
template <int MODE>
__device__ float end_height(const uint8_t *P, int sx, int sz, int rad = 12) {
    const int cx = sx / 2, cz = sz / 2, ox = sx % 2, oz = sz % 2;
    float best;
    if (!MODE) {
        uint32_t r2 = (uint32_t)sx * (uint32_t)sx + (uint32_t)sz * (uint32_t)sz;
        float len = (float)sqrt((double)(float)(int32_t)r2);
        best = clampf_ref(100.0f - len * 8.0f, -100.0f, 80.0f);
    } else best = -100.0f;
    for (int dx = -rad; dx <= rad; dx++)
        for (int dz = -rad; dz <= rad; dz++) {
            int64_t tx = (int64_t)cx + dx, tz = (int64_t)cz + dz;
            if (tx * tx + tz * tz <= 4096) continue;
            double nv = simplex2(P, (double)tx, (double)tz);
            bool island = MODE ? ((float)nv < -0.9f) : (nv < (double)-0.9f);
            if (!island) continue;
            float size = fmodf(fabsf((float)tx) * 3439.0f + fabsf((float)tz) * 147.0f, 13.0f) + 9.0f;
            float xd = (float)(ox - dx * 2), zd = (float)(oz - dz * 2);
            float h = 100.0f - (float)sqrt((double)(xd * xd + zd * zd)) * size;
            h = clampf_ref(h, -100.0f, 80.0f);
            best = best > h ? best : h;
        }
    return best;
}
template <int MODE>
__device__ int end_biome(const McbWorld *__restrict__ W, const uint8_t *P, int qx, int qz, int rad = 12) {
    const int cx = qx >> 2, cz = qz >> 2;
    if ((int64_t)cx * cx + (int64_t)cz * cz <= 4096) return W->end_ids[0];
    const int bx = (cx * 2 + 1) * 8, bz = (cz * 2 + 1) * 8;
    float hgt = end_height<MODE>(P, bx / 8, bz / 8, rad);
    double h;
    if (!MODE) h = ((double)hgt - 8.0) / 128.0;
    else {
        float dx = (float)(0 - bx), dz = (float)(0 - bz);
        float len = (float)sqrt((double)(dx * dx + 0.0f * 0.0f + dz * dz));
        float c = (clampf_ref(100.0f - len, -100.0f, 80.0f) - 8.0f) * 0.0078125f;
        float e = (hgt - 8.0f) / 128.0f;
        h = (double)(c > e ? c : e);
    }
    if (h > 0.25) return W->end_ids[1];
    if (h >= -0.0625) return W->end_ids[2];
    return h < -0.21875 ? W->end_ids[3] : W->end_ids[4];
}

// ================================================================================================
// LDS staging of a seed slot's permutation tables
// ================================================================================================
// doubled tables: octave o at lds[o*512], its 256 bytes twice in a row
__device__ __forceinline__ void stage_perm2(uint8_t *lds, const McbSlot *slot, int n_oct) {
    const uint4 *src = reinterpret_cast<const uint4 *>(slot->perm[0]);
    uint4 *dst = reinterpret_cast<uint4 *>(lds);
    const int n16 = n_oct * 16;
    for (int i = threadIdx.x; i < n16; i += blockDim.x) {
        const uint4 v = src[i];
        const int o = i >> 4, q = i & 15;
        dst[o * 32 + q] = v;
        dst[o * 32 + 16 + q] = v;
    }
}
__device__ __forceinline__ void stage_perm(uint8_t *lds, const McbSlot *slot, int n_oct) {
    const uint4 *src = reinterpret_cast<const uint4 *>(slot->perm[0]);
    uint4 *dst = reinterpret_cast<uint4 *>(lds);
    const int n16 = n_oct * 16;
    for (int i = threadIdx.x; i < n16; i += blockDim.x) dst[i] = src[i];
}

#define END_CAP 1024
template <int MODE>
__device__ float end_height_list(const int *ix, const int *iz, const float *isz, int n, int sx, int sz) {
    const int cx = sx / 2, cz = sz / 2, ox = sx % 2, oz = sz % 2;
    float best;
    if (!MODE) {
        uint32_t r2 = (uint32_t)sx * (uint32_t)sx + (uint32_t)sz * (uint32_t)sz;
        float len = (float)sqrt((double)(float)(int32_t)r2);
        best = clampf_ref(100.0f - len * 8.0f, -100.0f, 80.0f);
    } else best = -100.0f;
    for (int i = 0; i < n; i++) {
        const int dx = ix[i] - cx, dz = iz[i] - cz;
        if (dx < -12 || dx > 12 || dz < -12 || dz > 12) continue;
        float xd = (float)(ox - dx * 2), zd = (float)(oz - dz * 2);
        float h = 100.0f - (float)sqrt((double)(xd * xd + zd * zd)) * isz[i];
        h = clampf_ref(h, -100.0f, 80.0f);
        best = best > h ? best : h;
    }
    return best;
}
// The End's erosion value at block (bx, bz) from the height value of its section (mc_end.h:96-106)
template <int MODE>
__device__ __forceinline__ double end_erosion_from_height(int bx, int bz, float hgt) {
    if (!MODE) return ((double)hgt - 8.0) / 128.0;
    float dx = (float)(0 - bx), dz = (float)(0 - bz);
    float len = (float)sqrt((double)(dx * dx + 0.0f * 0.0f + dz * dz));
    float c = (clampf_ref(100.0f - len, -100.0f, 80.0f) - 8.0f) * 0.0078125f;
    float e = (hgt - 8.0f) / 128.0f;
    return (double)(c > e ? c : e);
}
template <int MODE>
__device__ int end_biome_from_height(const McbWorld *__restrict__ W, int cx, int cz, float hgt) {
    const double h = end_erosion_from_height<MODE>((cx * 2 + 1) * 8, (cz * 2 + 1) * 8, hgt);
    if (h > 0.25) return W->end_ids[1];
    if (h >= -0.0625) return W->end_ids[2];
    return h < -0.21875 ? W->end_ids[3] : W->end_ids[4];
}


// ================================================================================================
// BiomeManager.obfuscateSeed on the device: SHA-256 (FIPS 180-4) of the seed's 8 little-endian bytes, the
// first 8 digest bytes read little-endian (libmcgen/src/world.c:118-152); one 64-byte block.
// ================================================================================================
__device__ __forceinline__ uint32_t rotr32(uint32_t v, int k) { return (v >> k) | (v << (32 - k)); }
__device__ int64_t hashed_seed_dev(int64_t seed) {
    constexpr uint32_t K[64] = {
        0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5, 0xd807aa98,
        0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174, 0xe49b69c1, 0xefbe4786,
        0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da, 0x983e5152, 0xa831c66d, 0xb00327c8,
        0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967, 0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13,
        0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85, 0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819,
        0xd6990624, 0xf40e3585, 0x106aa070, 0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a,
        0x5b9cca4f, 0x682e6ff3, 0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7,
        0xc67178f2};
    const uint64_t u = (uint64_t)seed;
    uint32_t w[16];
    w[0] = __builtin_bswap32((uint32_t)u);          // message bytes b0..b3 big-endian
    w[1] = __builtin_bswap32((uint32_t)(u >> 32));
    w[2] = 0x80000000u;
#pragma unroll
    for (int i = 3; i < 15; i++) w[i] = 0;
    w[15] = 64;
    uint32_t a = 0x6a09e667, b = 0xbb67ae85, c = 0x3c6ef372, d = 0xa54ff53a, e = 0x510e527f, f = 0x9b05688c,
             g = 0x1f83d9ab, h = 0x5be0cd19;
#pragma unroll
    for (int i = 0; i < 64; i++) {
        if (i >= 16) {
            const uint32_t x15 = w[(i - 15) & 15], x2 = w[(i - 2) & 15];
            const uint32_t s0 = rotr32(x15, 7) ^ rotr32(x15, 18) ^ (x15 >> 3);
            const uint32_t s1 = rotr32(x2, 17) ^ rotr32(x2, 19) ^ (x2 >> 10);
            w[i & 15] = w[i & 15] + s0 + w[(i - 7) & 15] + s1;
        }
        const uint32_t t1 = h + (rotr32(e, 6) ^ rotr32(e, 11) ^ rotr32(e, 25)) + ((e & f) ^ (~e & g)) + K[i] + w[i & 15];
        const uint32_t t2 = (rotr32(a, 2) ^ rotr32(a, 13) ^ rotr32(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
        h = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
    }
    const uint32_t d0 = 0x6a09e667 + a, d1 = 0xbb67ae85 + b;   // digest words 0 and 1 (big-endian bytes)
    return (int64_t)((uint64_t)__builtin_bswap32(d0) | (uint64_t)__builtin_bswap32(d1) << 32);
}

// BiomeManager zoom steps (engine/cuda/mc_gpu.cuh:96-118)
__device__ __forceinline__ int64_t zstep(int64_t r, int64_t c) {
    uint64_t v = (uint64_t)r;
    return (int64_t)(v * (v * 6364136223846793005ULL + 1442695040888963407ULL) + (uint64_t)c);
}
__device__ __forceinline__ double zfiddle(int64_t r) { return ((double)((r >> 24) & 1023) / 1024.0 - 0.5) * 0.9; }

// Quart cell BiomeManager.getBiome(x, y, z) reads for block (x, y, z) under hashed seed hs: the lattice
// corner of least fiddled distance, first minimum in corner order (engine/cuda/mc_gpu.cuh:96-118).
// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2): block b runs on XCD
// b % 8.  xcd_rank ranks the blocks XCD by XCD, so consecutive ranks (tiles of one seed, neighbouring
// rows of one request) run on one XCD and share its L2; a bijection of [0, nwg).
__device__ __forceinline__ int xcd_rank(int b, int nwg) { return nwg % 8 == 0 ? (b % 8) * (nwg / 8) + b / 8 : b; }

// The three offsets of lattice point (cx, cy, cz): they depend only on the point and the hashed seed, so
// kernels over grids compute each once and share it between the blocks around the point.
struct ZFid {
    double x, y, z;
};
__device__ __forceinline__ ZFid zoom_fiddles(int64_t hs, int64_t cx, int64_t cy, int64_t cz) {
    int64_t r = hs;
    r = zstep(r, cx); r = zstep(r, cy); r = zstep(r, cz);
    r = zstep(r, cx); r = zstep(r, cy); r = zstep(r, cz);
    ZFid f;
    f.x = zfiddle(r);
    r = zstep(r, hs);
    f.y = zfiddle(r);
    r = zstep(r, hs);
    f.z = zfiddle(r);
    return f;
}
// The corner pick for block (x, y, z) with corner c's offsets from fid(c) (c = ux << 2 | uy << 1 | uz).
template <class Fid>
__device__ __forceinline__ void zoom_pick(int x, int y, int z, Fid fid, int *q) {
    const int ax = x - 2, ay = y - 2, az = z - 2;
    const int px = ax >> 2, py = ay >> 2, pz = az >> 2;
    const double fx = (double)(ax & 3) / 4.0, fy = (double)(ay & 3) / 4.0, fz = (double)(az & 3) / 4.0;
    int best = 0;
    double bd = __builtin_inf();
    for (int c = 0; c < 8; c++) {
        const int ux = (c >> 2) & 1, uy = (c >> 1) & 1, uz = c & 1;
        const double dx = ux ? fx - 1.0 : fx, dy = uy ? fy - 1.0 : fy, dz = uz ? fz - 1.0 : fz;
        const ZFid o = fid(c, px + ux, py + uy, pz + uz);
        const double d = (dz + o.z) * (dz + o.z) + (dy + o.y) * (dy + o.y) + (dx + o.x) * (dx + o.x);
        if (bd > d) { best = c; bd = d; }
    }
    q[0] = px + ((best >> 2) & 1);
    q[1] = py + ((best >> 1) & 1);
    q[2] = pz + (best & 1);
}
__device__ __forceinline__ void zoom_quart(int64_t hs, int x, int y, int z, int *q) {
    zoom_pick(x, y, z, [&](int, int cx, int cy, int cz) { return zoom_fiddles(hs, cx, cy, cz); }, q);
}

// Biomes of every leaf at the least distance (all ties), by a walk that enters a subtree when its
// bound is <= the best distance (engine/cuda/mc_gpu.cuh:166-187 semantics; mask bit = biome id).
__device__ void tree_tie_mask(const McbNode *__restrict__ N, int n, const int *t, uint64_t *m) {
    double bd = __builtin_inf();
    uint64_t m0 = 0, m1 = 0;
    int cur = 1;
    while (cur < n) {
        const McbNode nd = N[cur];
        const double cd = node_dist(nd, t);
        if (cd <= bd) {
            if (nd.leaf) {
                if (cd < bd) { bd = cd; m0 = m1 = 0; }
                const int b = nd.leaf & 0xFF;
                if (b < 64) m0 |= 1ULL << b; else m1 |= 1ULL << (b - 64);
                cur = nd.escape;
            } else {
                cur++;
            }
        } else {
            cur = nd.escape;
        }
    }
    m[0] = m0;
    m[1] = m1;
}

}  // namespace dev
}  // namespace mcb
#endif
