// mcg_plugin_dev.h - device pieces of the libmcgen GPU plugin shared by its kernels (mcg_plugin.hip:
// biome cells and points; mcg_terrain.hip: the terrain volume executor and ore veins): Java arithmetic,
// octaves and NoiseStacks of both branches, End islands, splines over value slots, and the point
// evaluators of 26.3+ (float) and 26.1/26.2 (double) programs.  Semantics and operation order are
// described where mcg_plugin.hip introduces them.
#ifndef MCG_PLUGIN_DEV_H
#define MCG_PLUGIN_DEV_H

#include "mcg_plugin.h"
#include "mcb_device.h"

namespace mcb {
namespace plug {

using dev::GradD;
using dev::GradEnt;
using dev::gdot_d;
using dev::gdot_f;

// dynamic LDS layout of k_cells: gradient table at 0, then the permutation tables (when staged),
// then the value slots
template <bool NEWF> struct GradOf { typedef GradEnt T; };
template <> struct GradOf<false> { typedef GradD T; };
template <bool NEWF> constexpr int grad_bytes() { return 256 * (int)sizeof(typename GradOf<NEWF>::T); }
static_assert(grad_bytes<true>() == grad_lds_bytes(true) && grad_bytes<false>() == grad_lds_bytes(false), "gradient table size");

// ---------------------------------------------------------------------------------------------------
// Java arithmetic (mcgpu_dev.cuh:24-58)
// ---------------------------------------------------------------------------------------------------
__device__ __forceinline__ int32_t jd2i(double v) {
    if (v != v) return 0;
    if (v >= 2147483647.0) return 2147483647;
    if (v <= -2147483648.0) return (int32_t)0x80000000u;
    return (int32_t)v;
}
__device__ __forceinline__ int64_t jd2l(double v) {
    if (v != v) return 0;
    if (v >= 9223372036854775807.0) return INT64_MAX;
    if (v <= -9223372036854775808.0) return INT64_MIN;
    return (int64_t)v;
}
__device__ __forceinline__ int32_t jfloor(double v) { return jd2i(floor(v)); }
__device__ __forceinline__ int32_t jfloordiv(int32_t a, int32_t b) {
    int32_t q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) q--;
    return q;
}
__device__ __forceinline__ int32_t jfloormod(int32_t a, int32_t b) {
    int32_t m = a % b;
    if (m != 0 && ((m < 0) != (b < 0))) m += b;
    return m;
}
__device__ __forceinline__ float jlerpf(float a, float p0, float p1) { return p0 + a * (p1 - p0); }
__device__ __forceinline__ double jlerp(double a, double p0, double p1) { return p0 + a * (p1 - p0); }
__device__ __forceinline__ float jmaxf(float a, float b) {
    if (a != a) return a;
    if (a == 0.0f && b == 0.0f && signbit(a)) return b;
    return a >= b ? a : b;
}
__device__ __forceinline__ float jminf(float a, float b) {
    if (a != a) return a;
    if (a == 0.0f && b == 0.0f && signbit(b)) return b;
    return a <= b ? a : b;
}
__device__ __forceinline__ double jmax(double a, double b) {
    if (a != a) return a;
    if (a == 0.0 && b == 0.0 && signbit(a)) return b;
    return a >= b ? a : b;
}
__device__ __forceinline__ double jmin(double a, double b) {
    if (a != a) return a;
    if (a == 0.0 && b == 0.0 && signbit(b)) return b;
    return a <= b ? a : b;
}
__device__ __forceinline__ float jclampf(float v, float lo, float hi) { return v < lo ? lo : jminf(v, hi); }
__device__ __forceinline__ double jclamp(double v, double lo, double hi) { return v < lo ? lo : jmin(v, hi); }
__device__ __forceinline__ float jsignf(float v) {
    if (v != v || v == 0.0f) return v;
    return v > 0.0f ? 1.0f : -1.0f;
}

// ---------------------------------------------------------------------------------------------------
// octaves (mcgpu_dev.cuh:64-155; CPU noise.c)
// ---------------------------------------------------------------------------------------------------
// an octave's offsets: the head of its McgOct record, or its copy in LDS
struct Off3 {
    double xo, yo, zo;
};
// Permutation lookup: D2 = the doubled 512-byte table (the coordinates are reduced mod 256 once, every
// later index stays below 512, so perm[i & 255] = P[i] without masks); else the 256-byte table, masked.
template <bool D2> __device__ __forceinline__ int pperm(const uint8_t *P, int i) { return D2 ? (int)P[i] : (int)P[i & 255]; }
template <bool D2> __device__ __forceinline__ int pcoord(int v) { return D2 ? (v & 255) : v; }
__device__ __forceinline__ float smooth_f(float x) { return x * x * x * (x * (x * 6.0f - 15.0f) + 10.0f); }
__device__ __forceinline__ double smooth_d(double x) { return x * x * x * (x * (x * 6.0 - 15.0) + 10.0); }
__device__ __forceinline__ double wrap_new(double x) {
    const double HALF = 16777215.999999998;
    return (x >= -HALF && x < HALF) ? x : x - floor(x / 3.3554432E7 + 0.5) * 3.3554432E7;
}
__device__ __forceinline__ double wrap_old(double x) { return x - (double)jd2l(floor(x / 3.3554432E7 + 0.5)) * 3.3554432E7; }

#define PP(i) pperm<D2>(P, i)
template <bool D2>
__device__ __forceinline__ float sample_lerp_f(const uint8_t *P, const GradEnt *G, int x, int y, int z, float rx, float ry,
                                               float rz, float ry_orig) {
    x = pcoord<D2>(x); y = pcoord<D2>(y); z = pcoord<D2>(z);
    const int x0 = PP(x), x1 = PP(x + 1);
    const int xy00 = PP(x0 + y), xy01 = PP(x0 + y + 1), xy10 = PP(x1 + y), xy11 = PP(x1 + y + 1);
    const float gx = rx - 1.0f, gy = ry - 1.0f, gz = rz - 1.0f;
    const float d000 = gdot_f(G[PP(xy00 + z)], rx, ry, rz), d100 = gdot_f(G[PP(xy10 + z)], gx, ry, rz);
    const float d010 = gdot_f(G[PP(xy01 + z)], rx, gy, rz), d110 = gdot_f(G[PP(xy11 + z)], gx, gy, rz);
    const float d001 = gdot_f(G[PP(xy00 + z + 1)], rx, ry, gz), d101 = gdot_f(G[PP(xy10 + z + 1)], gx, ry, gz);
    const float d011 = gdot_f(G[PP(xy01 + z + 1)], rx, gy, gz), d111 = gdot_f(G[PP(xy11 + z + 1)], gx, gy, gz);
    const float a1 = smooth_f(rx), a2 = smooth_f(ry_orig), a3 = smooth_f(rz);
    return jlerpf(a3, jlerpf(a2, jlerpf(a1, d000, d100), jlerpf(a1, d010, d110)),
                  jlerpf(a2, jlerpf(a1, d001, d101), jlerpf(a1, d011, d111)));
}
// GradientNoise (26.3+): perlin (kind 0) or smeared (kind 1, y fudge fs)
template <bool D2>
__device__ __forceinline__ float octave_f(const uint8_t *P, const GradEnt *G, const Off3 *o, int kind, double fs, double _x,
                                          double _y, double _z) {
    const double x = wrap_new(_x) + o->xo, y = wrap_new(_y) + o->yo, z = wrap_new(_z) + o->zo;
    const int32_t fx = jfloor(x), fy = jfloor(y), fz = jfloor(z);
    const float rx = (float)(x - fx), rz = (float)(z - fz);
    const double ry = y - fy;
    float fry = (float)ry;
    if (kind != 0) {
        const double lim = (_y >= 0.0 && _y < ry) ? _y : ry;
        fry = (float)(ry - (double)jfloor(lim / fs + (double)1.0E-7f) * fs);
    }
    return sample_lerp_f<D2>(P, G, fx, fy, fz, rx, fry, rz, (float)ry);
}
// ImprovedNoise.noise of the legacy path, coordinates already wrapped (yScale = yFudge = 0*factor)
template <bool D2>
__device__ __forceinline__ double octave_d(const uint8_t *P, const GradD *G, const Off3 *o, double _x, double _y, double _z,
                                           double y_scale, double y_fudge) {
    const double x = _x + o->xo, y = _y + o->yo, z = _z + o->zo;
    const int32_t xf = jfloor(x), yf = jfloor(y), zf = jfloor(z);
    const double xr = x - xf, yr = y - yf, zr = z - zf;
    double yrf = 0.0;
    if (y_scale != 0.0) {
        const double lim = (y_fudge >= 0.0 && y_fudge < yr) ? y_fudge : yr;
        yrf = (double)jfloor(lim / y_scale + (double)1.0E-7f) * y_scale;
    }
    const double yy = yr - yrf;
    const int xm = pcoord<D2>(xf), ym = pcoord<D2>(yf), zm = pcoord<D2>(zf);
    const int x0 = PP(xm), x1 = PP(xm + 1);
    const int xy00 = PP(x0 + ym), xy01 = PP(x0 + ym + 1), xy10 = PP(x1 + ym), xy11 = PP(x1 + ym + 1);
    const double d000 = gdot_d(G[PP(xy00 + zm)], xr, yy, zr), d100 = gdot_d(G[PP(xy10 + zm)], xr - 1.0, yy, zr);
    const double d010 = gdot_d(G[PP(xy01 + zm)], xr, yy - 1.0, zr), d110 = gdot_d(G[PP(xy11 + zm)], xr - 1.0, yy - 1.0, zr);
    const double d001 = gdot_d(G[PP(xy00 + zm + 1)], xr, yy, zr - 1.0), d101 = gdot_d(G[PP(xy10 + zm + 1)], xr - 1.0, yy, zr - 1.0);
    const double d011 = gdot_d(G[PP(xy01 + zm + 1)], xr, yy - 1.0, zr - 1.0);
    const double d111 = gdot_d(G[PP(xy11 + zm + 1)], xr - 1.0, yy - 1.0, zr - 1.0);
    const double a1 = smooth_d(xr), a2 = smooth_d(yr), a3 = smooth_d(zr);
    return jlerp(a3, jlerp(a2, jlerp(a1, d000, d100), jlerp(a1, d010, d110)),
                 jlerp(a2, jlerp(a1, d001, d101), jlerp(a1, d011, d111)));
}
#undef PP

// bytes per staged permutation table: doubled (no index masks) or single (masked, a smaller LDS footprint)
constexpr int kPermStride = kPermDoubled ? 512 : 256;
// lds = base of the dynamic LDS: gradient table, then (PL) the permutation tables and copies of
// the noise records, layer records and octave offsets; the octave loop reads its uniform records there (a
// broadcast LDS read) instead of through a chain of global loads.
template <bool NEWF, bool PL>
__device__ __forceinline__ const uint8_t *perm_of(const PWorld &W, const uint8_t *lds, int oct) {
    return PL ? lds + grad_bytes<NEWF>() + oct * kPermStride : W.octs[oct].p;
}
__device__ __forceinline__ int rec_base(const PWorld &W, int gbytes) { return gbytes + W.noct * kPermStride; }
__device__ __forceinline__ int al16(int v) { return (v + 15) & ~15; }
template <bool NEWF, bool PL>
__device__ __forceinline__ const McgNoise *noise_of(const PWorld &W, const uint8_t *lds, int i) {
    return PL ? (const McgNoise *)(lds + rec_base(W, grad_bytes<NEWF>())) + i : W.noise + i;
}
template <bool NEWF, bool PL>
__device__ __forceinline__ const McgLayer *layer_of(const PWorld &W, const uint8_t *lds, int i) {
    return PL ? (const McgLayer *)(lds + rec_base(W, grad_bytes<NEWF>()) + al16(W.nnoise * (int)sizeof(McgNoise))) + i
              : W.layers + i;
}
template <bool NEWF, bool PL>
__device__ __forceinline__ const Off3 *off_of(const PWorld &W, const uint8_t *lds, int oct) {
    return PL ? (const Off3 *)(lds + rec_base(W, grad_bytes<NEWF>()) + al16(W.nnoise * (int)sizeof(McgNoise)) +
                               al16(W.nlayers * (int)sizeof(McgLayer))) + oct
              : (const Off3 *)(W.octs + oct);
}
// bytes of the staged records after the permutation tables
__device__ __forceinline__ int rec_bytes(const PWorld &W) {
    return al16(W.nnoise * (int)sizeof(McgNoise)) + al16(W.nlayers * (int)sizeof(McgLayer)) + al16(W.noct * (int)sizeof(Off3));
}
// NoiseStack value (26.3+; mcgpu_dev.cuh:157-168, noise.c:339-347)
template <bool PL>
__device__ float ns_get(const PWorld &W, const uint8_t *lp, int nsi, double x, double y, double z) {
    const McgNoise mn = *noise_of<true, PL>(W, lp, nsi);
    float v = 0.0f;
#pragma unroll 1
    for (int i = 0; i < mn.n0; i++) {
        const McgLayer L = *layer_of<true, PL>(W, lp, mn.l0 + i);
        const double f = L.a;
        const float g = octave_f<PL && kPermDoubled>(perm_of<true, PL>(W, lp, L.oct), (const GradEnt *)lp, off_of<true, PL>(W, lp, L.oct), L.kind,
                                     L.b, x * f, y * f, z * f);
        v += L.amp_f * g;
    }
    return v;
}
// legacy PerlinNoise.getValue (yScale = yFudge = 0) over one layer range (mcgpu_dev.cuh:170-181)
template <bool PL>
__device__ double old_perlin(const PWorld &W, const uint8_t *lp, int l0, int n, double x, double y, double z) {
    double value = 0.0;
#pragma unroll 1
    for (int i = 0; i < n; i++) {
        const McgLayer L = *layer_of<false, PL>(W, lp, l0 + i);
        if (L.oct >= 0) {
            const double factor = L.a;
            const double nv = octave_d<PL && kPermDoubled>(perm_of<false, PL>(W, lp, L.oct), (const GradD *)lp, off_of<false, PL>(W, lp, L.oct),
                                           wrap_old(x * factor), wrap_old(y * factor),
                                       wrap_old(z * factor), 0.0 * factor, 0.0 * factor);
            value += L.amp_d * nv * L.b;
        }
    }
    return value;
}
template <bool PL>
__device__ double old_normal(const PWorld &W, const uint8_t *lp, int nsi, double x, double y, double z) {
    const McgNoise mn = *noise_of<false, PL>(W, lp, nsi);
    const double x2 = x * 1.0181268882175227, y2 = y * 1.0181268882175227, z2 = z * 1.0181268882175227;
    return (old_perlin<PL>(W, lp, mn.l0, mn.n0, x, y, z) + old_perlin<PL>(W, lp, mn.l1, mn.n1, x2, y2, z2)) *
           mn.value_factor;
}

// ---------------------------------------------------------------------------------------------------
// End islands height (mcgpu_dev.cuh:212-243; df_new.c:659-673, df_old.c:449-465) with Java's clamp and
// max: a NaN centre term (the 26.1/26.2 section square wraps negative beyond |section| ~ 32768) stays NaN
// ---------------------------------------------------------------------------------------------------
template <bool OLD>
__device__ float end_height_j(const uint8_t *P, int sx, int sz) {
    const int cx = sx / 2, cz = sz / 2, subx = sx % 2, subz = sz % 2;
    float doffs = -100.0f;
    if (OLD) {
        const int32_t sq = (int32_t)((uint32_t)sx * (uint32_t)sx + (uint32_t)sz * (uint32_t)sz);
        doffs = jclampf(100.0f - (float)sqrt((double)(float)sq) * 8.0f, -100.0f, 80.0f);
    }
#pragma unroll 1
    for (int xo = -12; xo <= 12; xo++)
#pragma unroll 1
        for (int zo = -12; zo <= 12; zo++) {
            const int64_t tx = (int64_t)cx + xo, tz = (int64_t)cz + zo;
            if (tx * tx + tz * tz <= 4096LL) continue;
            const double nv = dev::simplex2(P, (double)tx, (double)tz);
            if (OLD ? !(nv < (double)-0.9f) : !((float)nv < -0.9f)) continue;
            const float size = fmodf(fabsf((float)tx) * 3439.0f + fabsf((float)tz) * 147.0f, 13.0f) + 9.0f;
            const float xd = (float)(subx - xo * 2), zd = (float)(subz - zo * 2);
            const float nd = jclampf(100.0f - (float)sqrt((double)(xd * xd + zd * zd)) * size, -100.0f, 80.0f);
            doffs = jmaxf(doffs, nd);
        }
    return doffs;
}

// The islands node at block (bx, bz): the height of section (bx / 8, bz / 8) from the End table when that
// section is a table chunk's (section 2c + 1 of chunk c) and the node reads the table's octave, else the
// window computed here.
template <bool OLD>
__device__ __forceinline__ float end_height_at(const PWorld &W, const uint8_t *P, int oct, int bx, int bz) {
    const int sx = bx / 8, sz = bz / 8;
    if (W.etab && oct == W.etab_oct && (sx & 1) && (sz & 1)) {
        const int ix = ((sx - 1) >> 1) - W.etab_x0, iz = ((sz - 1) >> 1) - W.etab_z0;
        if ((unsigned)ix < (unsigned)W.etab_nx && (unsigned)iz < (unsigned)W.etab_nz) return W.etab[(int64_t)iz * W.etab_nx + ix];
    }
    return end_height_j<OLD>(P, sx, sz);
}

// ---------------------------------------------------------------------------------------------------
// spline (mcgpu_dev.cuh:246-289, df_new.c:707-728): bounded compile-time recursion, one call site per
// level; coordinates are the spline node's coordinate slots
// ---------------------------------------------------------------------------------------------------
template <typename VT, int D>
__device__ float spline(const PWorld &W, int spi, const VT *vals, int arr0) {
    if constexpr (D == 0) {
        return __builtin_nanf("");                      // unreachable: the host bounds the depth
    } else {
        const McgSpline s = W.sp[spi];
        if (s.is_const) return s.value;
        const float in = (float)vals[W.arr[arr0 + s.coord] * kWG];
        const float *loc = W.splf + s.loc0, *der = W.splf + s.loc0 + s.n;
        int from = 0, len = s.n;
        while (len > 0) {
            const int half = len / 2, mid = from + half;
            if (in < loc[mid]) len = half;
            else { from = mid + 1; len -= half + 1; }
        }
        const int start = from - 1, last = s.n - 1;
        const bool edge = start < 0 || start == last;
        const int base = start < 0 ? 0 : start;
        float ya = 0.0f, yb = 0.0f;
#pragma unroll 1
        for (int q = 0; q < (edge ? 1 : 2); q++) {
            const float v = spline<VT, D - 1>(W, W.spc[s.val0 + base + q], vals, arr0);
            if (q == 0) ya = v; else yb = v;
        }
        if (edge) {
            const float d = der[base];
            return d == 0.0f ? ya : ya + d * (in - loc[base]);
        }
        const float x1 = loc[start], x2 = loc[start + 1];
        const float t = (in - x1) / (x2 - x1);
        const float d1 = der[start], d2 = der[start + 1];
        const float a = d1 * (x2 - x1) - (yb - ya);
        const float b = -d2 * (x2 - x1) + (yb - ya);
        return jlerpf(t, ya, yb) + t * (1.0f - t) * jlerpf(t, a, b);
    }
}

// ---------------------------------------------------------------------------------------------------
// program evaluation: every node in order, values into the lane's LDS slots
// ---------------------------------------------------------------------------------------------------
__device__ __forceinline__ float squeeze_f(float v) {
    const float x = jclampf(v, -1.0f, 1.0f);
    return x / 2.0f - (x * x * x) / 24.0f;
}
__device__ __forceinline__ float grad_value(const PIns &s, int coord) {    // df_new.c:676-694
    if (s.op == MCGK_GRAD_CLAMP) {
        const int cc = coord < s.i2 ? s.i2 : (coord > s.i3 ? s.i3 : coord);
        return s.f0 + (float)(cc - s.i1) * s.f1;
    }
    if (s.op == MCGK_GRAD_REPEAT) return s.f0 + (float)jfloormod(coord - s.i1, s.i2) * s.f1;
    const int rel = coord - s.i1;
    const int tile = jfloordiv(rel, s.i2);
    const int local = rel - tile * s.i2;
    return (tile & 1) == 0 ? s.f0 + (float)local * s.f1 : s.f0 + (float)(s.i2 - local) * s.f1;
}
__device__ __forceinline__ float dist_value(int m, float dx, float dy, float dz) {         // df_new.c:695-702
    switch (m) {
    case 0: return (float)sqrt((double)(dx * dx + dy * dy + dz * dz));
    case 1: return dx * dx + dy * dy + dz * dz;
    case 2: return fabsf(dx) + fabsf(dy) + fabsf(dz);
    default: return jmaxf(jmaxf(fabsf(dx), fabsf(dy)), fabsf(dz));
    }
}

// Header fields of an instruction record (opcode, operand slots, coordinate overrides): wave-uniform,
// read through the scalar cache by way of the constant address space.  The rest of the record stays
// vector loads, which keeps the scalar register count within bounds.
struct PHdr { int32_t op, dst, a, b, c, cm, cx, cy, cz, ns; float f0; double d0, d1; };
typedef const __attribute__((address_space(4))) int32_t *const_i32p;
__device__ __forceinline__ PHdr phdr(const PIns *p) {
    const const_i32p q = (const_i32p)p;
    const __attribute__((address_space(4))) double *qd = (const __attribute__((address_space(4))) double *)p;
    return PHdr{q[0], q[1], q[2], q[3], q[4], q[15], q[16], q[17], q[18], q[5], __int_as_float(q[20]), qd[12], qd[13]};
}
static_assert(offsetof(PIns, cm) == 15 * 4 && offsetof(PIns, cz) == 18 * 4 && offsetof(PIns, ns) == 5 * 4 &&
                  offsetof(PIns, f0) == 20 * 4 && offsetof(PIns, d0) == 12 * 8 && offsetof(PIns, d1) == 13 * 8,
              "PIns header layout");

// 26.3+ float program (mcgpu_dev.cuh:305-374; df_new.c:740-836 with caches off)
template <bool PL, int SPD>
__device__ void eval_new(const PWorld &W, const uint8_t *lp, float *vals, int ox, int oy, int oz, bool same) {
#define V(x) vals[(x) * kWG]
    // a wave whose lanes all repeat their x and z runs only the y-dependent copies
    const bool all = __all(same);
    const int i0 = all ? W.nins : 0, i1 = all ? W.nins + W.nydep : W.nins;
#pragma unroll 1
    for (int i = i0; i < i1; i++) {
        const PIns s = W.ins[i];
        const PHdr h = phdr(W.ins + i);
        if (same && (h.cm & kCmYFree)) continue;           // value kept from this lane's evaluation at the same x, z
        const int bx = (h.cm & 1) ? h.cx : ox, by = (h.cm & 2) ? h.cy : oy, bz = (h.cm & 4) ? h.cz : oz;
        float r;
        switch (h.op) {
        case MCGK_CONST: r = h.f0; break;
        case MCGK_NOISE: r = ns_get<PL>(W, lp, h.ns, bx * h.d0, by * h.d1, bz * h.d0); break;
        case MCGK_NOISE_XZ: {
            const double nx = bx * h.d0 + (double)V(h.a), ny = by * h.d1, nz = bz * h.d0 + (double)V(h.c);
            r = ns_get<PL>(W, lp, h.ns, nx, ny, nz);
            break;
        }
        case MCGK_NOISE_XYZ: {
            const double nx = bx * h.d0 + (double)V(h.a), ny = by * h.d1 + (double)V(h.b), nz = bz * h.d0 + (double)V(h.c);
            r = ns_get<PL>(W, lp, h.ns, nx, ny, nz);
            break;
        }
        case MCGK_SHIFT_B: r = ns_get<PL>(W, lp, h.ns, bz * h.d0, bx * h.d0, 0.0) * 4.0f; break;
        case MCGK_END: r = (end_height_at<false>(W, perm_of<true, PL>(W, lp, s.oct), s.oct, bx, bz) - 8.0f) / 128.0f; break;
        case MCGK_DIST: r = dist_value(s.i3, (float)(s.i0 - bx), (float)(s.i1 - by), (float)(s.i2 - bz)); break;
        case MCGK_GRAD_CLAMP: case MCGK_GRAD_REPEAT: case MCGK_GRAD_MIRROR:
            r = grad_value(s, s.i0 == 0 ? bx : (s.i0 == 1 ? by : bz));
            break;
        case MCGK_CTX_ALPHA: case MCGK_CTX_OFFSET: case MCGK_CTX_BEARD: r = h.f0; break;
        case MCGK_ABS: r = fabsf(V(h.a)); break;
        case MCGK_SQUARE: { const float v = V(h.a); r = v * v; break; }
        case MCGK_CUBE: { const float v = V(h.a); r = v * v * v; break; }
        case MCGK_SQRT: r = (float)sqrt((double)V(h.a)); break;
        case MCGK_LEAKY: { const float v = V(h.a); r = v > 0.0f ? v : v * h.f0; break; }
        case MCGK_RECIP: r = 1.0f / V(h.a); break;
        case MCGK_NEG: r = -V(h.a); break;
        case MCGK_SQUEEZE: r = squeeze_f(V(h.a)); break;
        case MCGK_SIGN: r = jsignf(V(h.a)); break;
        case MCGK_ADD: r = V(h.a) + V(h.b); break;
        case MCGK_CADD: r = V(h.a) + h.f0; break;
        case MCGK_CSUB: r = h.f0 - V(h.a); break;
        case MCGK_SUB: r = V(h.a) - V(h.b); break;
        case MCGK_MUL: { const float l = V(h.a); r = l == 0.0f ? 0.0f : l * V(h.b); break; }
        case MCGK_CMUL: r = V(h.a) * h.f0; break;
        case MCGK_DIV: { const float l = V(h.a); r = l == 0.0f ? 0.0f : l / V(h.b); break; }
        case MCGK_CDIV: r = h.f0 / V(h.a); break;
        case MCGK_MIN: { const float l = V(h.a); r = l <= h.f0 ? l : jminf(l, V(h.b)); break; }
        case MCGK_CMIN: r = jminf(V(h.a), h.f0); break;
        case MCGK_MAX: { const float l = V(h.a); r = l >= h.f0 ? l : jmaxf(l, V(h.b)); break; }
        case MCGK_CMAX: r = jmaxf(V(h.a), h.f0); break;
        case MCGK_SPLINE: r = spline<float, SPD>(W, s.sp, vals, s.arr0); break;
        case MCGK_LERP: {
            const float a = V(h.a);
            r = a == 0.0f ? V(h.b) : (a == 1.0f ? V(h.c) : jlerpf(a, V(h.b), V(h.c)));
            break;
        }
        case MCGK_LERP_CF: {
            const float a = V(h.a);
            r = a == 0.0f ? h.f0 : (a == 1.0f ? V(h.c) : jlerpf(a, h.f0, V(h.c)));
            break;
        }
        case MCGK_LERP_CS: {
            const float a = V(h.a);
            r = a == 0.0f ? V(h.b) : (a == 1.0f ? h.f0 : jlerpf(a, V(h.b), h.f0));
            break;
        }
        case MCGK_CLAMP: r = jclampf(V(h.a), h.f0, s.f1); break;
        case MCGK_RANGE_C: { const float v = V(h.a); r = (v >= h.f0 && v < s.f1) ? s.f2 : s.f3; break; }
        case MCGK_RANGE: { const float v = V(h.a); r = (v >= h.f0 && v < s.f1) ? V(h.b) : V(h.c); break; }
        case MCGK_ISEL1: { const float v = V(h.a); r = v < h.f0 ? V(h.b) : V(h.c); break; }
        case MCGK_ISEL: {
            const float v = V(h.a);
            int k = 0;
            for (; k < s.nthr; k++)
                if (v < (float)W.thr[s.thr0 + k]) break;
            if (k == s.nthr) k = s.narr - 1;
            r = V(W.arr[s.arr0 + k]);
            break;
        }
        default: r = V(h.a); break;            // CACHE, BLEND_DENSITY, SLICE_*: the argument's value
        }
        V(h.dst) = r;
    }
#undef V
}

// 26.1/26.2 double program (mcgpu_dev.cuh:387-437; df_old.c point evaluation)
__device__ __forceinline__ double clamped_map(double v, double a, double b, double ta, double tb) {
    const double f = (v - a) / (b - a);
    return f < 0.0 ? ta : (f > 1.0 ? tb : jlerp(f, ta, tb));
}
__device__ __forceinline__ double weird_rarity(int type, double r) {
    if (type == 0) return r < -0.5 ? 0.75 : (r < 0.0 ? 1.0 : (r < 0.5 ? 1.5 : 2.0));
    return r < -0.75 ? 0.5 : (r < -0.5 ? 0.75 : (r < 0.5 ? 1.0 : (r < 0.75 ? 2.0 : 3.0)));
}
template <bool PL, int SPD>
__device__ void eval_old(const PWorld &W, const uint8_t *lp, double *vals, int bx, int by, int bz, bool same) {
#define V(x) vals[(x) * kWG]
    // a wave whose lanes all repeat their x and z runs only the y-dependent copies
    const bool all = __all(same);
    const int i0 = all ? W.nins : 0, i1 = all ? W.nins + W.nydep : W.nins;
#pragma unroll 1
    for (int i = i0; i < i1; i++) {
        const PIns s = W.ins[i];
        const PHdr h = phdr(W.ins + i);
        if (same && (h.cm & kCmYFree)) continue;           // value kept from this lane's evaluation at the same x, z
        double r;
        switch (h.op) {
        case MCGO_CONST: r = h.d0; break;
        case MCGO_NOISE: r = old_normal<PL>(W, lp, h.ns, bx * h.d0, by * h.d1, bz * h.d0); break;
        case MCGO_SHIFTED_NOISE: {
            const double x = bx * h.d0 + V(h.a), y = by * h.d1 + V(h.b), z = bz * h.d0 + V(h.c);
            r = old_normal<PL>(W, lp, h.ns, x, y, z);
            break;
        }
        case MCGO_SHIFT_A: r = old_normal<PL>(W, lp, h.ns, (double)bx * h.d0, 0.0 * h.d0, (double)bz * h.d0) * 4.0; break;
        case MCGO_SHIFT_B: r = old_normal<PL>(W, lp, h.ns, (double)bz * h.d0, (double)bx * h.d0, 0.0 * h.d0) * 4.0; break;
        case MCGO_SHIFT: r = old_normal<PL>(W, lp, h.ns, (double)bx * h.d0, (double)by * h.d1, (double)bz * h.d0) * 4.0; break;
        case MCGO_END_ISLANDS:
            r = ((double)end_height_at<true>(W, perm_of<false, PL>(W, lp, s.oct), s.oct, bx, bz) - 8.0) / 128.0;
            break;
        case MCGO_Y_GRADIENT: r = clamped_map((double)by, (double)s.i0, (double)s.i1, h.d0, h.d1); break;
        case MCGO_WEIRD_SCALED: {
            const double rr = weird_rarity(s.i0, V(h.a));
            r = rr * fabs(old_normal<PL>(W, lp, h.ns, bx / rr, by / rr, bz / rr));
            break;
        }
        case MCGO_BLEND_ALPHA: r = 1.0; break;
        case MCGO_ZERO: r = 0.0; break;
        case MCGO_ABS: r = fabs(V(h.a)); break;
        case MCGO_SQUARE: { const double v = V(h.a); r = v * v; break; }
        case MCGO_CUBE: { const double v = V(h.a); r = v * v * v; break; }
        case MCGO_HALF_NEG: { const double v = V(h.a); r = v > 0.0 ? v : v * 0.5; break; }
        case MCGO_QUARTER_NEG: { const double v = V(h.a); r = v > 0.0 ? v : v * 0.25; break; }
        case MCGO_RECIP: r = 1.0 / V(h.a); break;
        case MCGO_SQUEEZE: { const double c = jclamp(V(h.a), -1.0, 1.0); r = c / 2.0 - c * c * c / 24.0; break; }
        case MCGO_CLAMP: r = jclamp(V(h.a), h.d0, h.d1); break;
        case MCGO_ADD: r = V(h.a) + V(h.b); break;
        case MCGO_MUL: { const double v = V(h.a); r = v == 0.0 ? 0.0 : v * V(h.b); break; }
        case MCGO_MIN: { const double v = V(h.a); r = v < h.d1 ? v : jmin(v, V(h.b)); break; }
        case MCGO_MAX: { const double v = V(h.a); r = v > h.d1 ? v : jmax(v, V(h.b)); break; }
        case MCGO_CADD: r = V(h.a) + h.d0; break;
        case MCGO_CMUL: r = V(h.a) * h.d0; break;
        case MCGO_RANGE: { const double v = V(h.a); r = (v >= h.d0 && v < h.d1) ? V(h.b) : V(h.c); break; }
        case MCGO_ISEL: {
            const double v = V(h.a);
            int k = 0;
            for (; k < s.nthr; k++)
                if (v < W.thr[s.thr0 + k]) break;
            if (k == s.nthr) k = s.narr - 1;
            r = V(W.arr[s.arr0 + k]);
            break;
        }
        case MCGO_SPLINE: r = (double)spline<double, SPD>(W, s.sp, vals, s.arr0); break;
        default: r = V(h.a); break;            // PASS
        }
        V(h.dst) = r;
    }
#undef V
}

template <bool NEWF> struct VT_of;
template <> struct VT_of<true> { typedef float T; };
template <> struct VT_of<false> { typedef double T; };

template <bool NEWF, bool PL, int SPD>
__device__ __forceinline__ void eval_prog(const PWorld &W, const uint8_t *lp, typename VT_of<NEWF>::T *vals, int bx, int by,
                                          int bz, bool same) {
    if constexpr (NEWF) eval_new<PL, SPD>(W, lp, vals, bx, by, bz, same);
    else eval_old<PL, SPD>(W, lp, vals, bx, by, bz, same);
}

}  // namespace plug
}  // namespace mcb
#endif
