/* mcgplugin.c - CPU restatement of libmcgen's biome lookup on a McgBiomeWorld (the GPU plugin's input).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the plugin entry points of libmcbiome_hip.so (mcgpu_*).
 * Only tests/ load it.
 *
 * It follows libmcgen's CPU path, not the reference plugin: density functions are evaluated lazily and
 * recursively as libmcgen/src/df_new.c:740-836 (26.3+, float) and df_old.c:471-520 (26.1/26.2, double)
 * do, on the exported node graph; the tree search is Climate.RTree.search with a candidate
 * (engine/mc_biomes.h:31-62); the chunk-fill-order tie chain is world_biome_cell
 * (libmcgen/src/biome.c:64-120) without the plugin's length cap.  Beside each result it reports
 * whether the plugin rules leave the position unresolved (value 255): a non-finite or out-of-range
 * climate value, a chain of more than 12 tied cells, or no leaf found.
 *
 * Two formulations of the same semantics are here so tests can check one against the other:
 * or_mcg_cell (the backward chain) and or_mcg_chunk_seq (one sequential pass over a chunk's cells in
 * fill order, each search starting from the previous result, libmcgen/src/biome.c:143-150).
 */
#include "mcgplugin.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t i64;
typedef uint64_t u64;
typedef int32_t i32;

/* ---- Java arithmetic (libmcgen/src/mcgen_internal.h jm_*) ---- */
static i32 d2i(double v) {
    if (v != v) return 0;
    if (v >= 2147483647.0) return 2147483647;
    if (v <= -2147483648.0) return (i32)0x80000000u;
    return (i32)v;
}
static i64 d2l(double v) {
    if (v != v) return 0;
    if (v >= 9223372036854775807.0) return INT64_MAX;
    if (v <= -9223372036854775808.0) return INT64_MIN;
    return (i64)v;
}
static i32 ffloor(double v) { return d2i(floor(v)); }
static i32 floordiv(i32 a, i32 b) { i32 q = a / b; if ((a % b != 0) && ((a < 0) != (b < 0))) q--; return q; }
static i32 floormod(i32 a, i32 b) { i32 m = a % b; if (m != 0 && ((m < 0) != (b < 0))) m += b; return m; }
static float lerpf(float a, float p0, float p1) { return p0 + a * (p1 - p0); }
static double lerpd(double a, double p0, double p1) { return p0 + a * (p1 - p0); }
static float maxf_j(float a, float b) { if (a != a) return a; if (a == 0.0f && b == 0.0f && signbit(a)) return b; return a >= b ? a : b; }
static float minf_j(float a, float b) { if (a != a) return a; if (a == 0.0f && b == 0.0f && signbit(b)) return b; return a <= b ? a : b; }
static double max_j(double a, double b) { if (a != a) return a; if (a == 0.0 && b == 0.0 && signbit(a)) return b; return a >= b ? a : b; }
static double min_j(double a, double b) { if (a != a) return a; if (a == 0.0 && b == 0.0 && signbit(b)) return b; return a <= b ? a : b; }
static float clampf_j(float v, float lo, float hi) { return v < lo ? lo : minf_j(v, hi); }
static double clamp_j(double v, double lo, double hi) { return v < lo ? lo : min_j(v, hi); }

/* ---- octaves (libmcgen/src/noise.c:55-127) ---- */
static const signed char G3[16][3] = {
    {1,1,0},{-1,1,0},{1,-1,0},{-1,-1,0},{1,0,1},{-1,0,1},{1,0,-1},{-1,0,-1},
    {0,1,1},{0,-1,1},{0,1,-1},{0,-1,-1},{1,1,0},{0,-1,1},{-1,1,0},{0,-1,-1} };
#define PERM(o, i) ((int)(o)->p[(i) & 255])
static float dotf(int h, float x, float y, float z) { const signed char *g = G3[h & 15]; return (float)g[0] * x + (float)g[1] * y + (float)g[2] * z; }
static double dotd(int h, double x, double y, double z) { const signed char *g = G3[h & 15]; return g[0] * x + g[1] * y + g[2] * z; }
static float smoothf(float x) { return x * x * x * (x * (x * 6.0f - 15.0f) + 10.0f); }
static double smoothd(double x) { return x * x * x * (x * (x * 6.0 - 15.0) + 10.0); }
static double wrap_new(double x) {
    const double HALF = 16777215.999999998;
    return (x >= -HALF && x < HALF) ? x : x - floor(x / 3.3554432E7 + 0.5) * 3.3554432E7;
}
static double wrap_old(double x) { return x - (double)d2l(floor(x / 3.3554432E7 + 0.5)) * 3.3554432E7; }

static float lerp_corners_f(const McgOct *o, int x, int y, int z, float rx, float ry, float rz, float ry_orig) {
    int x0 = PERM(o, x), x1 = PERM(o, x + 1);
    int a = PERM(o, x0 + y), b = PERM(o, x0 + y + 1), c = PERM(o, x1 + y), d = PERM(o, x1 + y + 1);
    float v000 = dotf(PERM(o, a + z), rx, ry, rz), v100 = dotf(PERM(o, c + z), rx - 1.0f, ry, rz);
    float v010 = dotf(PERM(o, b + z), rx, ry - 1.0f, rz), v110 = dotf(PERM(o, d + z), rx - 1.0f, ry - 1.0f, rz);
    float v001 = dotf(PERM(o, a + z + 1), rx, ry, rz - 1.0f), v101 = dotf(PERM(o, c + z + 1), rx - 1.0f, ry, rz - 1.0f);
    float v011 = dotf(PERM(o, b + z + 1), rx, ry - 1.0f, rz - 1.0f), v111 = dotf(PERM(o, d + z + 1), rx - 1.0f, ry - 1.0f, rz - 1.0f);
    float u = smoothf(rx), v = smoothf(ry_orig), w = smoothf(rz);
    return lerpf(w, lerpf(v, lerpf(u, v000, v100), lerpf(u, v010, v110)), lerpf(v, lerpf(u, v001, v101), lerpf(u, v011, v111)));
}
static float perlin_f(const McgOct *o, double X, double Y, double Z) {                   /* noise.c:83-88 */
    double x = wrap_new(X) + o->xo, y = wrap_new(Y) + o->yo, z = wrap_new(Z) + o->zo;
    i32 fx = ffloor(x), fy = ffloor(y), fz = ffloor(z);
    float rx = (float)(x - fx), ry = (float)(y - fy), rz = (float)(z - fz);
    return lerp_corners_f(o, fx, fy, fz, rx, ry, rz, ry);
}
static float smeared_f(const McgOct *o, double fs, double X, double Y, double Z) {      /* noise.c:90-103 */
    double x = wrap_new(X) + o->xo, y = wrap_new(Y) + o->yo, z = wrap_new(Z) + o->zo;
    i32 fx = ffloor(x), fy = ffloor(y), fz = ffloor(z);
    double ry = y - fy;
    double lim = (Y >= 0.0 && Y < ry) ? Y : ry;
    double fudge = (double)ffloor(lim / fs + (double)1.0E-7f) * fs;
    return lerp_corners_f(o, fx, fy, fz, (float)(x - fx), (float)(ry - fudge), (float)(z - fz), (float)ry);
}
static double improved_d(const McgOct *o, double X, double Y, double Z, double ys, double yf) {   /* noise.c:105-125 */
    double x = X + o->xo, y = Y + o->yo, z = Z + o->zo;
    i32 xi = ffloor(x), yi = ffloor(y), zi = ffloor(z);
    double xr = x - xi, yr = y - yi, zr = z - zi, yrf = 0.0;
    if (ys != 0.0) {
        double lim = (yf >= 0.0 && yf < yr) ? yf : yr;
        yrf = (double)ffloor(lim / ys + (double)1.0E-7f) * ys;
    }
    double yy = yr - yrf;
    int x0 = PERM(o, xi), x1 = PERM(o, xi + 1);
    int a = PERM(o, x0 + yi), b = PERM(o, x0 + yi + 1), c = PERM(o, x1 + yi), d = PERM(o, x1 + yi + 1);
    double v000 = dotd(PERM(o, a + zi), xr, yy, zr), v100 = dotd(PERM(o, c + zi), xr - 1.0, yy, zr);
    double v010 = dotd(PERM(o, b + zi), xr, yy - 1.0, zr), v110 = dotd(PERM(o, d + zi), xr - 1.0, yy - 1.0, zr);
    double v001 = dotd(PERM(o, a + zi + 1), xr, yy, zr - 1.0), v101 = dotd(PERM(o, c + zi + 1), xr - 1.0, yy, zr - 1.0);
    double v011 = dotd(PERM(o, b + zi + 1), xr, yy - 1.0, zr - 1.0), v111 = dotd(PERM(o, d + zi + 1), xr - 1.0, yy - 1.0, zr - 1.0);
    double u = smoothd(xr), v = smoothd(yr), w = smoothd(zr);
    return lerpd(w, lerpd(v, lerpd(u, v000, v100), lerpd(u, v010, v110)), lerpd(v, lerpd(u, v001, v101), lerpd(u, v011, v111)));
}
static float stack_f(const McgProg *P, int ns, double x, double y, double z) {          /* noise.c:339-347 */
    const McgNoise *mn = &P->noise[ns];
    float v = 0.0f;
    for (int i = 0; i < mn->n0; i++) {
        const McgLayer *L = &P->layers[mn->l0 + i];
        const McgOct *o = &P->octs[L->oct];
        double f = L->a;
        float g = L->kind == 0 ? perlin_f(o, x * f, y * f, z * f) : smeared_f(o, L->b, x * f, y * f, z * f);
        v += L->amp_f * g;
    }
    return v;
}
static double perlin_old(const McgProg *P, int l0, int n, double x, double y, double z) {      /* noise.c:578-588 */
    double value = 0.0;
    for (int i = 0; i < n; i++) {
        const McgLayer *L = &P->layers[l0 + i];
        if (L->oct < 0) continue;
        double factor = L->a;
        double nv = improved_d(&P->octs[L->oct], wrap_old(x * factor), wrap_old(y * factor), wrap_old(z * factor), 0.0 * factor, 0.0 * factor);
        value += L->amp_d * nv * L->b;
    }
    return value;
}
static double normal_old(const McgProg *P, int ns, double x, double y, double z) {            /* noise.c:600-603 */
    const McgNoise *mn = &P->noise[ns];
    double x2 = x * 1.0181268882175227, y2 = y * 1.0181268882175227, z2 = z * 1.0181268882175227;
    return (perlin_old(P, mn->l0, mn->n0, x, y, z) + perlin_old(P, mn->l1, mn->n1, x2, y2, z2)) * mn->value_factor;
}

/* ---- End islands (df_new.c:659-673, df_old.c:449-465; simplex: engine/mc_end.h:38-80) ---- */
static double simplex2(const McgOct *o, double xin, double yin) {
    const double SQ3 = 1.7320508075688772, F2 = 0.5 * (SQ3 - 1.0), G2 = (3.0 - SQ3) / 6.0;
    double s = (xin + yin) * F2;
    i32 i = ffloor(xin + s), j = ffloor(yin + s);
    double t = (i + j) * G2, x0 = xin - (i - t), y0 = yin - (j - t);
    int i1 = x0 > y0 ? 1 : 0, j1 = 1 - i1;
    double x1 = x0 - i1 + G2, y1 = y0 - j1 + G2, x2 = x0 - 1.0 + 2.0 * G2, y2 = y0 - 1.0 + 2.0 * G2;
    int ii = i & 255, jj = j & 255;
    int g[3] = { PERM(o, ii + PERM(o, jj)) % 12, PERM(o, ii + i1 + PERM(o, jj + j1)) % 12, PERM(o, ii + 1 + PERM(o, jj + 1)) % 12 };
    double xs[3] = { x0, x1, x2 }, ys[3] = { y0, y1, y2 }, n = 0.0;
    double c[3];
    for (int k = 0; k < 3; k++) {
        double t0 = 0.5 - xs[k] * xs[k] - ys[k] * ys[k] - 0.0 * 0.0;
        if (t0 < 0.0) { c[k] = 0.0; continue; }
        t0 *= t0;
        c[k] = t0 * t0 * (G3[g[k]][0] * xs[k] + G3[g[k]][1] * ys[k] + G3[g[k]][2] * 0.0);
    }
    n = c[0] + c[1] + c[2];
    return 70.0 * n;
}
static float end_height(const McgOct *o, int sx, int sz, int old) {
    int cx = sx / 2, cz = sz / 2, subx = sx % 2, subz = sz % 2;
    float doffs;
    if (old) {
        i32 sq = (i32)((uint32_t)sx * (uint32_t)sx + (uint32_t)sz * (uint32_t)sz);
        doffs = clampf_j(100.0f - (float)sqrt((double)(float)sq) * 8.0f, -100.0f, 80.0f);
    } else doffs = -100.0f;
    for (int xo = -12; xo <= 12; xo++)
        for (int zo = -12; zo <= 12; zo++) {
            i64 tx = (i64)cx + xo, tz = (i64)cz + zo;
            if (tx * tx + tz * tz <= 4096LL) continue;
            double nv = simplex2(o, (double)tx, (double)tz);
            if (old ? !(nv < (double)-0.9f) : !((float)nv < -0.9f)) continue;
            float size = fmodf(fabsf((float)tx) * 3439.0f + fabsf((float)tz) * 147.0f, 13.0f) + 9.0f;
            float xd = (float)(subx - xo * 2), zd = (float)(subz - zo * 2);
            float nd = clampf_j(100.0f - (float)sqrt((double)(xd * xd + zd * zd)) * size, -100.0f, 80.0f);
            doffs = maxf_j(doffs, nd);
        }
    return doffs;
}

/* ---- spline (df_new.c:707-728, df_old.c:417-430): coordinates evaluated lazily, once per call ---- */
typedef struct { const McgProg *P; const McgNode *s; int x, y, z; double cv[64]; char have[64]; } SplCtx;
static double vnode(const McgProg *P, int i, int x, int y, int z);
static float spl_coord(SplCtx *c, int ci) {
    if (!c->have[ci]) { c->cv[ci] = vnode(c->P, c->P->arr[c->s->arr0 + ci], c->x, c->y, c->z); c->have[ci] = 1; }
    return (float)c->cv[ci];
}
static float spl(SplCtx *c, int spi) {
    const McgSpline *s = &c->P->sp[spi];
    if (s->is_const) return s->value;
    const float *loc = c->P->splf + s->loc0, *der = loc + s->n;
    const int *ch = c->P->spc + s->val0;
    float in = spl_coord(c, s->coord);
    int from = 0, len = s->n;
    while (len > 0) { int half = len / 2, mid = from + half; if (in < loc[mid]) len = half; else { from = mid + 1; len -= half + 1; } }
    int start = from - 1, last = s->n - 1;
    if (start < 0 || start == last) {
        int e = start < 0 ? 0 : last;
        float v = spl(c, ch[e]);
        return der[e] == 0.0f ? v : v + der[e] * (in - loc[e]);
    }
    float x1 = loc[start], x2 = loc[start + 1], t = (in - x1) / (x2 - x1);
    float y1 = spl(c, ch[start]), y2 = spl(c, ch[start + 1]);
    float a = der[start] * (x2 - x1) - (y2 - y1), b = -der[start + 1] * (x2 - x1) + (y2 - y1);
    return lerpf(t, y1, y2) + t * (1.0f - t) * lerpf(t, a, b);
}

/* ---- density functions, lazily ---- */
static float squeeze_f(float v) { float x = clampf_j(v, -1.0f, 1.0f); return x / 2.0f - (x * x * x) / 24.0f; }
static float grad_f(const McgNode *s, int coord) {                                          /* df_new.c:674-694 */
    if (s->k == MCGK_GRAD_CLAMP) { int cc = coord < s->i2 ? s->i2 : (coord > s->i3 ? s->i3 : coord); return s->f0 + (float)(cc - s->i1) * s->f1; }
    if (s->k == MCGK_GRAD_REPEAT) return s->f0 + (float)floormod(coord - s->i1, s->i2) * s->f1;
    int rel = coord - s->i1, tile = floordiv(rel, s->i2), local = rel - tile * s->i2;
    return (tile & 1) == 0 ? s->f0 + (float)local * s->f1 : s->f0 + (float)(s->i2 - local) * s->f1;
}
static float dist_f(int m, float dx, float dy, float dz) {                                  /* df_new.c:695-702 */
    if (m == 0) return (float)sqrt((double)(dx * dx + dy * dy + dz * dz));
    if (m == 1) return dx * dx + dy * dy + dz * dz;
    if (m == 2) return fabsf(dx) + fabsf(dy) + fabsf(dz);
    return maxf_j(maxf_j(fabsf(dx), fabsf(dy)), fabsf(dz));
}
static float vnew(const McgProg *P, int i, int ox, int oy, int oz) {                          /* df_new.c:740-836 */
    const McgNode *s = &P->nodes[i];
    const int bx = (s->cm & 1) ? s->cx : ox, by = (s->cm & 2) ? s->cy : oy, bz = (s->cm & 4) ? s->cz : oz;
#define A vnew(P, s->a, ox, oy, oz)
#define B vnew(P, s->b, ox, oy, oz)
#define C vnew(P, s->c, ox, oy, oz)
    switch (s->k) {
    case MCGK_CONST: return s->f0;
    case MCGK_NOISE: return stack_f(P, s->ns, bx * s->d0, by * s->d1, bz * s->d0);
    case MCGK_NOISE_XZ: { double nx = bx * s->d0 + (double)A, ny = by * s->d1, nz = bz * s->d0 + (double)C; return stack_f(P, s->ns, nx, ny, nz); }
    case MCGK_NOISE_XYZ: { double nx = bx * s->d0 + (double)A, ny = by * s->d1 + (double)B, nz = bz * s->d0 + (double)C; return stack_f(P, s->ns, nx, ny, nz); }
    case MCGK_SHIFT_B: return stack_f(P, s->ns, bz * s->d0, bx * s->d0, 0.0) * 4.0f;
    case MCGK_END: return (end_height(&P->octs[s->oct], bx / 8, bz / 8, 0) - 8.0f) / 128.0f;
    case MCGK_DIST: return dist_f(s->i3, (float)(s->i0 - bx), (float)(s->i1 - by), (float)(s->i2 - bz));
    case MCGK_GRAD_CLAMP: case MCGK_GRAD_REPEAT: case MCGK_GRAD_MIRROR: return grad_f(s, s->i0 == 0 ? bx : (s->i0 == 1 ? by : bz));
    case MCGK_CTX_ALPHA: case MCGK_CTX_OFFSET: case MCGK_CTX_BEARD: return s->f0;
    case MCGK_ABS: return fabsf(A);
    case MCGK_SQUARE: { float v = A; return v * v; }
    case MCGK_CUBE: { float v = A; return v * v * v; }
    case MCGK_SQRT: return (float)sqrt((double)A);
    case MCGK_LEAKY: { float v = A; return v > 0.0f ? v : v * s->f0; }
    case MCGK_RECIP: return 1.0f / A;
    case MCGK_NEG: return -A;
    case MCGK_SQUEEZE: return squeeze_f(A);
    case MCGK_SIGN: { float v = A; return (v != v || v == 0.0f) ? v : (v > 0.0f ? 1.0f : -1.0f); }
    case MCGK_ADD: { float l = A; return l + B; }
    case MCGK_CADD: return A + s->f0;
    case MCGK_CSUB: return s->f0 - A;
    case MCGK_SUB: { float l = A; return l - B; }
    case MCGK_MUL: { float l = A; return l == 0.0f ? 0.0f : l * B; }
    case MCGK_CMUL: return A * s->f0;
    case MCGK_DIV: { float l = A; return l == 0.0f ? 0.0f : l / B; }
    case MCGK_CDIV: return s->f0 / A;
    case MCGK_MIN: { float l = A; return l <= s->f0 ? l : minf_j(l, B); }
    case MCGK_CMIN: return minf_j(A, s->f0);
    case MCGK_MAX: { float l = A; return l >= s->f0 ? l : maxf_j(l, B); }
    case MCGK_CMAX: return maxf_j(A, s->f0);
    case MCGK_SPLINE: { SplCtx c; memset(&c, 0, sizeof c); c.P = P; c.s = s; c.x = ox; c.y = oy; c.z = oz; return spl(&c, s->sp); }
    case MCGK_LERP: { float a = A; if (a == 0.0f) return B; if (a == 1.0f) return C; float f = B; return lerpf(a, f, C); }
    case MCGK_LERP_CF: { float a = A; if (a == 0.0f) return s->f0; if (a == 1.0f) return C; return lerpf(a, s->f0, C); }
    case MCGK_LERP_CS: { float a = A; if (a == 0.0f) return B; if (a == 1.0f) return s->f0; return lerpf(a, B, s->f0); }
    case MCGK_CLAMP: return clampf_j(A, s->f0, s->f1);
    case MCGK_RANGE_C: { float v = A; return (v >= s->f0 && v < s->f1) ? s->f2 : s->f3; }
    case MCGK_RANGE: { float v = A; return (v >= s->f0 && v < s->f1) ? B : C; }
    case MCGK_ISEL1: { float v = A; return v < s->f0 ? B : C; }
    case MCGK_ISEL: {
        float v = A;
        int k = 0; for (; k < s->nthr; k++) if (v < (float)P->thr[s->thr0 + k]) break;
        if (k == s->nthr) k = s->narr - 1;
        return vnew(P, P->arr[s->arr0 + k], ox, oy, oz);
    }
    default: return A;              /* CACHE, BLEND_DENSITY, SLICE_* */
    }
#undef A
#undef B
#undef C
}
static double weird(int type, double r) {
    if (type == 0) return r < -0.5 ? 0.75 : (r < 0.0 ? 1.0 : (r < 0.5 ? 1.5 : 2.0));
    return r < -0.75 ? 0.5 : (r < -0.5 ? 0.75 : (r < 0.5 ? 1.0 : (r < 0.75 ? 2.0 : 3.0)));
}
static double vold(const McgProg *P, int i, int bx, int by, int bz) {                          /* df_old.c:471-520 */
    const McgNode *s = &P->nodes[i];
#define A vold(P, s->a, bx, by, bz)
#define B vold(P, s->b, bx, by, bz)
#define C vold(P, s->c, bx, by, bz)
    switch (s->k) {
    case MCGO_CONST: return s->d0;
    case MCGO_NOISE: return normal_old(P, s->ns, bx * s->d0, by * s->d1, bz * s->d0);
    case MCGO_SHIFTED_NOISE: { double x = bx * s->d0 + A; double y = by * s->d1 + B; double z = bz * s->d0 + C; return normal_old(P, s->ns, x, y, z); }
    case MCGO_SHIFT_A: return normal_old(P, s->ns, (double)bx * s->d0, 0.0 * s->d0, (double)bz * s->d0) * 4.0;
    case MCGO_SHIFT_B: return normal_old(P, s->ns, (double)bz * s->d0, (double)bx * s->d0, 0.0 * s->d0) * 4.0;
    case MCGO_SHIFT: return normal_old(P, s->ns, (double)bx * s->d0, (double)by * s->d1, (double)bz * s->d0) * 4.0;
    case MCGO_END_ISLANDS: return ((double)end_height(&P->octs[s->oct], bx / 8, bz / 8, 1) - 8.0) / 128.0;
    case MCGO_Y_GRADIENT: {
        double f = ((double)by - (double)s->i0) / ((double)s->i1 - (double)s->i0);
        return f < 0.0 ? s->d0 : (f > 1.0 ? s->d1 : lerpd(f, s->d0, s->d1));
    }
    case MCGO_WEIRD_SCALED: { double r = weird(s->i0, A); return r * fabs(normal_old(P, s->ns, bx / r, by / r, bz / r)); }
    case MCGO_BLEND_ALPHA: return 1.0;
    case MCGO_ZERO: return 0.0;
    case MCGO_ABS: return fabs(A);
    case MCGO_SQUARE: { double v = A; return v * v; }
    case MCGO_CUBE: { double v = A; return v * v * v; }
    case MCGO_HALF_NEG: { double v = A; return v > 0.0 ? v : v * 0.5; }
    case MCGO_QUARTER_NEG: { double v = A; return v > 0.0 ? v : v * 0.25; }
    case MCGO_RECIP: return 1.0 / A;
    case MCGO_SQUEEZE: { double c = clamp_j(A, -1.0, 1.0); return c / 2.0 - c * c * c / 24.0; }
    case MCGO_CLAMP: return clamp_j(A, s->d0, s->d1);
    case MCGO_ADD: { double v = A; return v + B; }
    case MCGO_MUL: { double v = A; return v == 0.0 ? 0.0 : v * B; }
    case MCGO_MIN: { double v = A; return v < s->d1 ? v : min_j(v, B); }
    case MCGO_MAX: { double v = A; return v > s->d1 ? v : max_j(v, B); }
    case MCGO_CADD: return A + s->d0;
    case MCGO_CMUL: return A * s->d0;
    case MCGO_RANGE: { double v = A; return (v >= s->d0 && v < s->d1) ? B : C; }
    case MCGO_ISEL: {
        double v = A;
        for (int k = 0; k < s->nthr; k++) if (v < P->thr[s->thr0 + k]) return vold(P, P->arr[s->arr0 + k], bx, by, bz);
        return vold(P, P->arr[s->arr0 + s->narr - 1], bx, by, bz);
    }
    case MCGO_SPLINE: { SplCtx c; memset(&c, 0, sizeof c); c.P = P; c.s = s; c.x = bx; c.y = by; c.z = bz; return (double)spl(&c, s->sp); }
    default: return A;              /* PASS */
    }
#undef A
#undef B
#undef C
}
static double vnode(const McgProg *P, int i, int x, int y, int z) { return P->old ? vold(P, i, x, y, z) : (double)vnew(P, i, x, y, z); }

float or_mcg_point_f(const McgProg *P, int node, int x, int y, int z) { return vnew(P, node, x, y, z); }
float or_mcg_stack_f(const McgProg *P, int ns, double x, double y, double z) { return stack_f(P, ns, x, y, z); }

int or_mcg_climate(const McgBiomeWorld *W, int bx, int by, int bz, double out6[6]) {
    const McgProg *P = W->prog;
    for (int k = 0; k < 6; k++) out6[k] = P->root[k] >= 0 ? vnode(P, P->root[k], bx, by, bz) : 0.0;
    return 0;
}

/* ---- R-tree: Climate.RTree.Node.search with a candidate (engine/mc_biomes.h:31-62) on the BFS array ---- */
static i64 ndist(const McgTreeNode *n, const i64 *t) {
    i64 d = 0;
    for (int k = 0; k < 7; k++) {
        i64 above = t[k] - (i64)n->hi[k], below = (i64)n->lo[k] - t[k];
        i64 x = above > 0 ? above : (below > 0 ? below : 0);
        d += x * x;
    }
    return d;
}
static int rt_search(const McgTreeNode *T, int ni, const i64 *t, int cand) {
    i64 min = cand < 0 ? INT64_MAX : ndist(&T[cand], t);
    int res = cand;
    for (int k = 0; k < T[ni].count; k++) {
        int ci = T[ni].first + k;
        i64 d = ndist(&T[ci], t);
        if (min > d) {
            int leaf = T[ci].count == 0 ? ci : rt_search(T, ci, t, res);
            i64 ld = leaf == ci ? d : ndist(&T[leaf], t);
            if (min > ld) { min = ld; res = leaf; }
        }
    }
    return res;
}
static int rt_count_min(const McgTreeNode *T, int ni, const i64 *t, i64 d) {               /* biome.c:73-82 */
    if (T[ni].count == 0) return ndist(&T[ni], t) == d;
    int c = 0;
    for (int k = 0; k < T[ni].count && c < 2; k++) {
        int ci = T[ni].first + k;
        if (ndist(&T[ci], t) <= d) c += rt_count_min(T, ci, t, d);
    }
    return c;
}

/* target of one quart cell; returns 0 when the plugin rules reject the climate (non-finite / huge) */
static int cell_tg(const McgBiomeWorld *W, int qx, int qy, int qz, i64 t[7]) {
    double v[6];
    or_mcg_climate(W, qx * 4, qy * 4, qz * 4, v);
    int ok = 1;
    for (int k = 0; k < 6; k++) {
        float f = (float)v[k] * 10000.0f;
        if (!(fabsf(f) < 9.0e18f)) ok = 0;
        t[k] = ok ? (i64)f : 0;
    }
    t[6] = 0;
    return ok;
}
static int clampq(const McgBiomeWorld *W, int qy) {
    return qy < W->min_qy ? W->min_qy : (qy > W->min_qy + W->qh - 1 ? W->min_qy + W->qh - 1 : qy);
}
static int end_cell(const McgBiomeWorld *W, int qx, int qy, int qz) {                      /* biome.c:26-45 */
    int cx = (qx * 4) >> 4, cz = (qz * 4) >> 4;
    if ((i64)cx * cx + (i64)cz * cz <= 4096LL) return W->end_ids[0];
    double v[6];
    or_mcg_climate(W, (cx * 2 + 1) * 8, qy * 4, (cz * 2 + 1) * 8, v);
    double h = v[3];
    if (h > 0.25) return W->end_ids[1];
    if (h >= -0.0625) return W->end_ids[2];
    return h < -0.21875 ? W->end_ids[3] : W->end_ids[4];
}

/* cap >= 0: stop once the chain is longer than cap (the plugin then reports the cell unresolved) */
static int cell_impl(const McgBiomeWorld *W, int qx, int qy, int qz, int *unres, int *chain, int cap);
int or_mcg_cell(const McgBiomeWorld *W, int qx, int qy, int qz, int *unres, int *chain) {  /* biome.c:87-120 */
    return cell_impl(W, qx, qy, qz, unres, chain, -1);
}
static int cell_impl(const McgBiomeWorld *W, int qx, int qy, int qz, int *unres, int *chain, int cap) {
    qy = clampq(W, qy);
    if (unres) *unres = 0;
    if (chain) *chain = 0;
    if (W->source == MCG_BS_FIXED) return W->fixed_biome;
    if (W->source == MCG_BS_END) return end_cell(W, qx, qy, qz);
    const McgTreeNode *T = W->tree;
    int bx = qx & ~3, bz = qz & ~3, ly = qy - W->min_qy;
    int idx = (ly >> 2) * 64 + (qx & 3) * 16 + (ly & 3) * 4 + (qz & 3);
    const int scap = idx + 1;                      /* the chain stops at fill index 0 at the latest */
    i64 (*st)[7] = malloc(sizeof(i64[7]) * (size_t)scap);
    i64 *sd = malloc(sizeof(i64) * (size_t)scap);
    int *sl = malloc(sizeof(int) * (size_t)scap);
    int n = 0, base = -1, bad = 0;
    for (;;) {
        int s = idx >> 6, r = idx & 63;
        i64 t[7];
        if (!cell_tg(W, bx + (r >> 4), W->min_qy + s * 4 + ((r >> 2) & 3), bz + (r & 3), t)) { bad = 1; break; }
        int l0 = rt_search(T, W->tree_root, t, -1);
        if (l0 < 0) { bad = 1; break; }
        i64 d = ndist(&T[l0], t);
        if (idx == 0 || rt_count_min(T, W->tree_root, t, d) < 2) { base = l0; break; }
        memcpy(st[n], t, sizeof t); sd[n] = d; sl[n] = l0; n++; idx--;
        if (cap >= 0 && n > cap) { bad = 1; break; }
    }
    int leaf = bad ? -1 : base;
    for (int i = n - 1; i >= 0 && leaf >= 0; i--) leaf = ndist(&T[leaf], st[i]) == sd[i] ? leaf : sl[i];
    free(st); free(sd); free(sl);
    if (chain) *chain = n;
    if (unres) *unres = bad || n > 12;
    return leaf >= 0 ? T[leaf].biome : -1;
}

int or_mcg_chunk_seq(const McgBiomeWorld *W, int cx, int cz, int32_t *out, int32_t *valid) {   /* biome.c:143-150 */
    const McgTreeNode *T = W->tree;
    int qh = W->qh, last = -1;
    for (int sy = 0; sy < qh; sy += 4)
        for (int qx = 0; qx < 4; qx++)
            for (int qy = sy; qy < sy + 4 && qy < qh; qy++)
                for (int qz = 0; qz < 4; qz++) {
                    i64 t[7];
                    int ok = cell_tg(W, cx * 4 + qx, W->min_qy + qy, cz * 4 + qz, t);
                    int leaf = rt_search(T, W->tree_root, t, last);
                    last = leaf;
                    out[(qy * 4 + qz) * 4 + qx] = leaf >= 0 ? T[leaf].biome : -1;
                    valid[(qy * 4 + qz) * 4 + qx] = ok;
                }
    return 0;
}

int or_mcg_points(const McgBiomeWorld *W, int n, const int *xyz, uint8_t *out, int *nunres, int cap) {
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : bad)
    for (int i = 0; i < n; i++) {
        int32_t q[3];
        or_zoom(W->zoom_seed, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], q);
        int u = 0;
        int b = cell_impl(W, q[0], q[1], q[2], &u, NULL, cap);
        if (u || b < 0) { out[i] = 255; bad++; } else out[i] = (uint8_t)b;
    }
    if (nunres) *nunres = bad;
    return 0;
}

int or_mcg_grid(const McgBiomeWorld *W, int x0, int z0, int nx, int nz, int step, int y, uint8_t *out, int *nunres, int cap) {
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : bad)
    for (int i = 0; i < nx * nz; i++) {
        int32_t q[3];
        or_zoom(W->zoom_seed, x0 + (i % nx) * step, y, z0 + (i / nx) * step, q);
        int u = 0;
        int b = cell_impl(W, q[0], q[1], q[2], &u, NULL, cap);
        if (u || b < 0) { out[i] = 255; bad++; } else out[i] = (uint8_t)b;
    }
    if (nunres) *nunres = bad;
    return 0;
}

/* Quart grid out[iz*nx+ix] at (qx0+ix, qy, qz0+iz), searched in row-major order, each search starting
 * from the previous sample's leaf (first from none): the order the reference's Java oracle grids use
 * (tests/vectors/<V>/biome_grid_hashes.tsv "point" rows).  Targets in parallel, searches in order. */
int or_mcg_quart_history(const McgBiomeWorld *W, int qx0, int qz0, int nx, int nz, int qy, uint8_t *out) {
    const int n = nx * nz;
    i64 (*t)[7] = malloc(sizeof(i64[7]) * (size_t)n);
    int ok_all = 1;
#pragma omp parallel for schedule(dynamic, 256) reduction(& : ok_all)
    for (int i = 0; i < n; i++) ok_all &= cell_tg(W, qx0 + i % nx, qy, qz0 + i / nx, t[i]);
    int last = -1;
    for (int i = 0; i < n; i++) {
        last = rt_search(W->tree, W->tree_root, t[i], last);
        out[i] = (uint8_t)W->tree[last].biome;
    }
    free(t);
    return ok_all ? 0 : -1;
}
