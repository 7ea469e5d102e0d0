/* mcgterrain.c - CPU restatement of libmcgen's 26.3+ terrain density path, on the program form a libmcgen
 * host hands the GPU plugin (McgProg, include/mcgpu.h).  TEST INFRASTRUCTURE ONLY: the product never links
 * it; tests check the HIP terrain executor against it.
 *
 * sampleVolume (libmcgen/src/df_new.c:917-1150) node kind by node kind, with the sampler context's cache
 * cells (df_new.c:611-636, 1073-1083) and spline coordinates evaluated lazily per volume
 * (df_new.c:705-728): the volume NoiseStack order (noise.c:141-190, 359-366: x and z gradient terms summed
 * first, coordinates block * (scale * frequency), amplitude amp * layer amplitude), the transposed
 * volume of shift_b, element-wise arithmetic as the volume path writes it (plain min/max comparisons, no
 * zero short-cut in mul/div), the cell-step interpolation (df_new.c:867-915) and the slices.  Point
 * pieces (NoiseStack.get of shifted noises, End islands, distance, gradients) come from mcgplugin.c.
 * Ore veins restate the GPU plugin's rule evaluation (libmcgen/gpu/mcgpu_terrain.cu:303-324) with the
 * ore PositionalRandomFactory (Mth.getSeed(x, y, z) ^ lo, hi; engine/mc_rng.h:80-131).
 * Node kinds outside the GPU executor's set (log, pow, round, find_top_surface) are refused. */
#include "mcgterrain.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mcgplugin.h"

typedef int32_t i32;
typedef int64_t i64;
typedef uint64_t u64;

typedef struct { int sx, sy, sz, x0, y0, z0, dx, dy, dz; } TV;
static int tv_n(const TV *v) { return v->sx * v->sy * v->sz; }
static int tv_idx(const TV *v, int ix, int iy, int iz) { return iy + (ix + iz * v->sx) * v->sy; }
static int tv_bx(const TV *v, int i) { return v->x0 + i * v->dx; }
static int tv_by(const TV *v, int i) { return v->y0 + i * v->dy; }
static int tv_bz(const TV *v, int i) { return v->z0 + i * v->dz; }
static int tv_eq(const TV *a, const TV *b) { return !memcmp(a, b, sizeof *a); }

static i32 d2i(double v) {
    if (v != v) return 0;
    if (v >= 2147483647.0) return 2147483647;
    if (v <= -2147483648.0) return (i32)0x80000000u;
    return (i32)v;
}
static i32 ffloor(double v) { return d2i(floor(v)); }
static i32 floordiv(i32 a, i32 b) { i32 q = a / b; if ((a % b != 0) && ((a < 0) != (b < 0))) q--; return q; }
static i32 floormod(i32 a, i32 b) { i32 m = a % b; if (m != 0 && ((m < 0) != (b < 0))) m += b; return m; }
static float lerpf(float a, float p0, float p1) { return p0 + a * (p1 - p0); }
static float smoothf(float x) { return x * x * x * (x * (x * 6.0f - 15.0f) + 10.0f); }
static float minf_j(float a, float b) { if (a != a) return a; if (a == 0.0f && b == 0.0f && signbit(b)) return b; return a <= b ? a : b; }
static float clampf_j(float v, float lo, float hi) { return v < lo ? lo : minf_j(v, hi); }
static float signf_j(float v) { if (v != v || v == 0.0f) return v; return v > 0.0f ? 1.0f : -1.0f; }
static float squeeze_f(float v) { float x = clampf_j(v, -1.0f, 1.0f); return x / 2.0f - (x * x * x) / 24.0f; }
static double wrap_new(double x) {
    const double HALF = 16777215.999999998;
    return (x >= -HALF && x < HALF) ? x : x - floor(x / 3.3554432E7 + 0.5) * 3.3554432E7;
}

/* the improved-noise gradient vectors, indexed by hash & 15 */
static const signed char GV[16][3] = {
    {1,1,0},{-1,1,0},{1,-1,0},{-1,-1,0},{1,0,1},{-1,0,1},{1,0,-1},{-1,0,-1},
    {0,1,1},{0,-1,1},{0,1,-1},{0,-1,-1},{1,1,0},{0,-1,1},{-1,1,0},{0,-1,-1} };
#define PM(o, i) ((int)(o)->p[(i) & 255])

/* PerlinNoise / SmearedPerlinNoise.addToVolume of one octave (noise.c:141-190) */
static void octave_add_volume(const McgOct *o, int smeared, double fs, float *buf, const TV *v, double sxz, double sy, float amp) {
    int index = 0;
    for (int iz = 0; iz < v->sz; iz++) {
        const double z = wrap_new(tv_bz(v, iz) * sxz) + o->zo;
        const i32 fz = ffloor(z);
        const float rz = (float)(z - fz), az = smoothf(rz);
        for (int ix = 0; ix < v->sx; ix++) {
            const double x = wrap_new(tv_bx(v, ix) * sxz) + o->xo;
            const i32 fx = ffloor(x);
            const float rx = (float)(x - fx), ax = smoothf(rx);
            const int x0 = PM(o, fx), x1 = PM(o, fx + 1);
            for (int iy = 0; iy < v->sy; iy++) {
                const double oy = tv_by(v, iy) * sy;
                const double y = wrap_new(oy) + o->yo;
                const i32 fy = ffloor(y);
                const double ryd = y - fy;
                const float ay = smoothf((float)ryd);
                float ry = smeared ? 0.0f : (float)ryd;
                const int h[8] = { PM(o, PM(o, x0 + fy) + fz), PM(o, PM(o, x1 + fy) + fz), PM(o, PM(o, x0 + fy + 1) + fz),
                                   PM(o, PM(o, x1 + fy + 1) + fz), PM(o, PM(o, x0 + fy) + fz + 1), PM(o, PM(o, x1 + fy) + fz + 1),
                                   PM(o, PM(o, x0 + fy + 1) + fz + 1), PM(o, PM(o, x1 + fy + 1) + fz + 1) };
                if (smeared) {
                    const double lim = (oy >= 0.0 && oy < ryd) ? oy : ryd;
                    ry = (float)(ryd - (double)ffloor(lim / fs + (double)1.0E-7f) * fs);
                }
                float c[8];
                for (int k = 0; k < 8; k++) {
                    const signed char *g = GV[h[k] & 15];
                    const float gx = (k & 1) ? rx - 1.0f : rx, gy = (k & 2) ? ry - 1.0f : ry, gz = (k & 4) ? rz - 1.0f : rz;
                    const float dxz = (float)g[0] * gx + (float)g[2] * gz;
                    c[k] = dxz + (float)g[1] * gy;
                }
                buf[index++] += amp * lerpf(az, lerpf(ay, lerpf(ax, c[0], c[1]), lerpf(ax, c[2], c[3])),
                                                lerpf(ay, lerpf(ax, c[4], c[5]), lerpf(ax, c[6], c[7])));
            }
        }
    }
}
/* NoiseStack.addToVolume (noise.c:359-366) */
static void stack_add_volume(const McgProg *P, int ns, float *buf, const TV *v, double sxz, double sy, float amp) {
    const McgNoise *mn = &P->noise[ns];
    for (int i = 0; i < mn->n0; i++) {
        const McgLayer *L = &P->layers[mn->l0 + i];
        const double f = L->a;
        octave_add_volume(&P->octs[L->oct], L->kind, L->b, buf, v, sxz * f, sy * f, amp * L->amp_f);
    }
}

typedef struct { int set; TV v; float *buf; int cap; } TCell;
typedef struct {
    const McgProg *P;
    TCell *cell; int ncell;
    McgBeardFn beard; void *ud; int chunk;
    int bad;
} TCtx;

static void vol(TCtx *x, int ni, float *o, const TV *v);

static float *acq(int n) { return (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1)); }

/* spline over coordinate buffers, each evaluated over the whole volume the first time an element needs it */
typedef struct { TCtx *x; const McgNode *s; const TV *v; float **bufs; int idx; } SplIn;
static float spl_coord(SplIn *in, int ci) {
    if (!in->bufs[ci]) {
        in->bufs[ci] = acq(tv_n(in->v));
        vol(in->x, in->x->P->arr[in->s->arr0 + ci], in->bufs[ci], in->v);
    }
    return in->bufs[ci][in->idx];
}
static float spl_eval(SplIn *in, int spi) {
    const McgProg *P = in->x->P;
    const McgSpline *s = &P->sp[spi];
    if (s->is_const) return s->value;
    const float *loc = P->splf + s->loc0, *der = loc + s->n;
    const int *ch = P->spc + s->val0;
    const float input = spl_coord(in, s->coord);
    int from = 0, len = s->n;
    while (len > 0) { int half = len / 2, mid = from + half; if (input < loc[mid]) len = half; else { from = mid + 1; len -= half + 1; } }
    const int start = from - 1, last = s->n - 1;
    if (start < 0 || start == last) {
        const int e = start < 0 ? 0 : last;
        const float v = spl_eval(in, ch[e]);
        return der[e] == 0.0f ? v : v + der[e] * (input - loc[e]);
    }
    const float x1 = loc[start], x2 = loc[start + 1], t = (input - x1) / (x2 - x1);
    const float d1 = der[start], d2 = der[start + 1];
    const float y1 = spl_eval(in, ch[start]), y2 = spl_eval(in, ch[start + 1]);
    const float a = d1 * (x2 - x1) - (y2 - y1), b = -d2 * (x2 - x1) + (y2 - y1);
    return lerpf(t, y1, y2) + t * (1.0f - t) * lerpf(t, a, b);
}

/* interpolation cell step (df_new.c:867-915) over a unit-step volume v */
static void interp_block(TCtx *x, const McgNode *s, float *out, const TV *v) {
    const int cxz = s->i0, cy = s->i1;
    const int mnx = floordiv(v->x0, cxz), mny = floordiv(v->y0, cy), mnz = floordiv(v->z0, cxz);
    const int maxx = v->x0 + v->sx * v->dx - 1, maxy = v->y0 + v->sy * v->dy - 1, maxz = v->z0 + v->sz * v->dz - 1;
    const int ncx = floordiv(maxx, cxz) - mnx + 1, ncy = floordiv(maxy, cy) - mny + 1, ncz = floordiv(maxz, cxz) - mnz + 1;
    const TV cv = { floormod(maxx, cxz) == 0 ? ncx : ncx + 1, floormod(maxy, cy) == 0 ? ncy : ncy + 1,
                    floormod(maxz, cxz) == 0 ? ncz : ncz + 1, mnx * cxz, mny * cy, mnz * cxz, cxz, cy, cxz };
    float *cb = acq(tv_n(&cv));
    vol(x, s->a, cb, &cv);
    for (int cz = 0; cz < ncz; cz++) {
        const int cz2 = cz + 1 < cv.sz - 1 ? cz + 1 : cv.sz - 1;
        for (int cxi = 0; cxi < ncx; cxi++) {
            const int cx2 = cxi + 1 < cv.sx - 1 ? cxi + 1 : cv.sx - 1;
            for (int cyi = 0; cyi < ncy; cyi++) {
                const int cy2 = cyi + 1 < cv.sy - 1 ? cyi + 1 : cv.sy - 1;
                const float v000 = cb[tv_idx(&cv, cxi, cyi, cz)], v100 = cb[tv_idx(&cv, cx2, cyi, cz)];
                const float v001 = cb[tv_idx(&cv, cxi, cyi, cz2)], v101 = cb[tv_idx(&cv, cx2, cyi, cz2)];
                const float v010 = cb[tv_idx(&cv, cxi, cy2, cz)], v110 = cb[tv_idx(&cv, cx2, cy2, cz)];
                const float v011 = cb[tv_idx(&cv, cxi, cy2, cz2)], v111 = cb[tv_idx(&cv, cx2, cy2, cz2)];
                /* the cell's blocks inside v */
                const int cox = tv_bx(&cv, cxi) - v->x0, coy = tv_by(&cv, cyi) - v->y0, coz = tv_bz(&cv, cz) - v->z0;
                const int xa = cox < 0 ? -cox : 0, ya = coy < 0 ? -coy : 0, za = coz < 0 ? -coz : 0;
                const int xb = (cxz < v->sx - cox ? cxz : v->sx - cox) - 1;
                const int yb = (cy < v->sy - coy ? cy : v->sy - coy) - 1;
                const int zb = (cxz < v->sz - coz ? cxz : v->sz - coz) - 1;
                for (int z = za; z <= zb; z++) {
                    const float az = (float)z * s->f0;
                    const float v00 = lerpf(az, v000, v001), v01 = lerpf(az, v010, v011);
                    const float v10 = lerpf(az, v100, v101), v11 = lerpf(az, v110, v111);
                    for (int xx = xa; xx <= xb; xx++) {
                        const float ax = (float)xx * s->f0;
                        const float a0 = lerpf(ax, v00, v10), a1 = lerpf(ax, v01, v11);
                        const float step = (a1 - a0) * s->f1;
                        float value = a0 + step * (float)ya;
                        int oi = tv_idx(v, cox + xx, coy + ya, coz + z);
                        for (int y = ya; y <= yb; y++) { out[oi++] = value; value += step; }
                    }
                }
            }
        }
    }
    free(cb);
}

static void fill(float *o, int n, float v) { for (int i = 0; i < n; i++) o[i] = v; }

static void vol(TCtx *x, int ni, float *o, const TV *v) {
    const McgProg *P = x->P;
    const McgNode *s = &P->nodes[ni];
    const int n = tv_n(v);
    switch (s->k) {
    case MCGK_CONST: case MCGK_CTX_ALPHA: case MCGK_CTX_OFFSET: fill(o, n, s->f0); return;
    case MCGK_CTX_BEARD: {
        const int vv[9] = { v->sx, v->sy, v->sz, v->x0, v->y0, v->z0, v->dx, v->dy, v->dz };
        if (!x->beard || !x->beard(x->ud, x->chunk, vv, o)) fill(o, n, s->f0);
        return;
    }
    case MCGK_NOISE: fill(o, n, 0.0f); stack_add_volume(P, s->ns, o, v, s->d0, s->d1, 1.0f); return;
    case MCGK_NOISE_XZ: case MCGK_NOISE_XYZ: {
        vol(x, s->a, o, v);
        float *by = NULL, *bz = acq(n);
        if (s->k == MCGK_NOISE_XYZ) { by = acq(n); vol(x, s->b, by, v); }
        vol(x, s->c, bz, v);
        int idx = 0;
        for (int iz = 0; iz < v->sz; iz++) {
            const double baseZ = tv_bz(v, iz) * s->d0;
            for (int ix = 0; ix < v->sx; ix++) {
                const double baseX = tv_bx(v, ix) * s->d0;
                for (int iy = 0; iy < v->sy; iy++, idx++) {
                    const double nx = baseX + (double)o[idx];
                    double ny = tv_by(v, iy) * s->d1;
                    if (by) ny += (double)by[idx];
                    const double nz = baseZ + (double)bz[idx];
                    o[idx] = or_mcg_stack_f(P, s->ns, nx, ny, nz);
                }
            }
        }
        free(bz); free(by);
        return;
    }
    case MCGK_SHIFT_B: {
        const TV tv = { v->sz, v->sx, 1, v->z0, v->x0, 0, v->dz, v->dx, 1 };
        float *tb = acq(tv_n(&tv));
        fill(tb, tv_n(&tv), 0.0f);
        stack_add_volume(P, s->ns, tb, &tv, s->d0, s->d0, 4.0f);
        for (int iz = 0; iz < v->sz; iz++)
            for (int ix = 0; ix < v->sx; ix++) fill(o + tv_idx(v, ix, 0, iz), v->sy, tb[tv_idx(&tv, iz, ix, 0)]);
        free(tb);
        return;
    }
    case MCGK_END:
        for (int iz = 0; iz < v->sz; iz++)
            for (int ix = 0; ix < v->sx; ix++) fill(o + tv_idx(v, ix, 0, iz), v->sy, or_mcg_point_f(P, ni, tv_bx(v, ix), 0, tv_bz(v, iz)));
        return;
    case MCGK_DIST: case MCGK_GRAD_CLAMP: case MCGK_GRAD_REPEAT: case MCGK_GRAD_MIRROR: {
        int idx = 0;
        for (int iz = 0; iz < v->sz; iz++)
            for (int ix = 0; ix < v->sx; ix++)
                for (int iy = 0; iy < v->sy; iy++) o[idx++] = or_mcg_point_f(P, ni, tv_bx(v, ix), tv_by(v, iy), tv_bz(v, iz));
        return;
    }
    case MCGK_ABS: vol(x, s->a, o, v); for (int i = 0; i < n; i++) o[i] = fabsf(o[i]); return;
    case MCGK_SQUARE: vol(x, s->a, o, v); for (int i = 0; i < n; i++) o[i] = o[i] * o[i]; return;
    case MCGK_CUBE: vol(x, s->a, o, v); for (int i = 0; i < n; i++) o[i] = o[i] * o[i] * o[i]; return;
    case MCGK_SQRT: vol(x, s->a, o, v); for (int i = 0; i < n; i++) o[i] = (float)sqrt((double)o[i]); return;
    case MCGK_LEAKY: vol(x, s->a, o, v); for (int i = 0; i < n; i++) o[i] = o[i] > 0.0f ? o[i] : o[i] * s->f0; return;
    case MCGK_RECIP: vol(x, s->a, o, v); for (int i = 0; i < n; i++) o[i] = 1.0f / o[i]; return;
    case MCGK_NEG: vol(x, s->a, o, v); for (int i = 0; i < n; i++) o[i] = -o[i]; return;
    case MCGK_SQUEEZE: vol(x, s->a, o, v); for (int i = 0; i < n; i++) o[i] = squeeze_f(o[i]); return;
    case MCGK_SIGN: vol(x, s->a, o, v); for (int i = 0; i < n; i++) o[i] = signf_j(o[i]); return;
    case MCGK_ADD: case MCGK_SUB: case MCGK_MUL: case MCGK_DIV: case MCGK_MIN: case MCGK_MAX: {
        vol(x, s->a, o, v);
        float *r = acq(n);
        vol(x, s->b, r, v);
        for (int i = 0; i < n; i++) {
            switch (s->k) {
            case MCGK_ADD: o[i] = o[i] + r[i]; break;
            case MCGK_SUB: o[i] = o[i] + -r[i]; break;
            case MCGK_MUL: o[i] = o[i] * r[i]; break;
            case MCGK_DIV: o[i] = o[i] / r[i]; break;
            case MCGK_MIN: if (r[i] < o[i]) o[i] = r[i]; break;
            default: if (r[i] > o[i]) o[i] = r[i]; break;
            }
        }
        free(r);
        return;
    }
    case MCGK_CADD: vol(x, s->a, o, v); for (int i = 0; i < n; i++) o[i] = o[i] + s->f0; return;
    case MCGK_CSUB: vol(x, s->a, o, v); for (int i = 0; i < n; i++) o[i] = s->f0 - o[i]; return;
    case MCGK_CMUL: vol(x, s->a, o, v); for (int i = 0; i < n; i++) o[i] = o[i] * s->f0; return;
    case MCGK_CDIV: vol(x, s->a, o, v); for (int i = 0; i < n; i++) o[i] = s->f0 / o[i]; return;
    case MCGK_CMIN: vol(x, s->a, o, v); for (int i = 0; i < n; i++) if (s->f0 < o[i]) o[i] = s->f0; return;
    case MCGK_CMAX: vol(x, s->a, o, v); for (int i = 0; i < n; i++) if (s->f0 > o[i]) o[i] = s->f0; return;
    case MCGK_SPLINE: {
        float *bufs[64] = { 0 };
        if (s->narr > 64) { x->bad = 1; return; }
        SplIn in = { x, s, v, bufs, 0 };
        for (int i = 0; i < n; i++) { in.idx = i; o[i] = spl_eval(&in, s->sp); }
        for (int k = 0; k < s->narr; k++) free(bufs[k]);
        return;
    }
    case MCGK_LERP: case MCGK_LERP_CF: case MCGK_LERP_CS: {
        vol(x, s->a, o, v);
        float *f = NULL, *sc = NULL;
        if (s->k != MCGK_LERP_CF) { f = acq(n); vol(x, s->b, f, v); }
        if (s->k != MCGK_LERP_CS) { sc = acq(n); vol(x, s->c, sc, v); }
        for (int i = 0; i < n; i++) {
            const float a = o[i], fv = f ? f[i] : s->f0, sv = sc ? sc[i] : s->f0;
            o[i] = a == 0.0f ? fv : (a == 1.0f ? sv : lerpf(a, fv, sv));
        }
        free(f); free(sc);
        return;
    }
    case MCGK_CLAMP: vol(x, s->a, o, v); for (int i = 0; i < n; i++) o[i] = clampf_j(o[i], s->f0, s->f1); return;
    case MCGK_RANGE_C: vol(x, s->a, o, v); for (int i = 0; i < n; i++) o[i] = (o[i] >= s->f0 && o[i] < s->f1) ? s->f2 : s->f3; return;
    case MCGK_RANGE: {
        vol(x, s->b, o, v);
        float *in = acq(n), *out = acq(n);
        vol(x, s->a, in, v);
        vol(x, s->c, out, v);
        for (int i = 0; i < n; i++) if (!(in[i] >= s->f0) || !(in[i] < s->f1)) o[i] = out[i];
        free(in); free(out);
        return;
    }
    case MCGK_ISEL1: {
        vol(x, s->a, o, v);
        float *bl = acq(n), *ab = acq(n);
        vol(x, s->b, bl, v);
        vol(x, s->c, ab, v);
        for (int i = 0; i < n; i++) o[i] = o[i] < s->f0 ? bl[i] : ab[i];
        free(bl); free(ab);
        return;
    }
    case MCGK_ISEL: {
        vol(x, s->a, o, v);
        float **bs = (float **)calloc((size_t)s->narr, sizeof(float *));
        for (int k = 0; k < s->narr; k++) { bs[k] = acq(n); vol(x, P->arr[s->arr0 + k], bs[k], v); }
        for (int i = 0; i < n; i++) {
            int k = 0;
            for (; k < s->nthr; k++) if (o[i] < (float)P->thr[s->thr0 + k]) break;
            if (k == s->nthr) k = s->narr - 1;
            o[i] = bs[k][i];
        }
        for (int k = 0; k < s->narr; k++) free(bs[k]);
        free(bs);
        return;
    }
    case MCGK_CACHE: {
        if (s->cid < 0 || s->cid >= x->ncell) { vol(x, s->a, o, v); return; }
        TCell *c = &x->cell[s->cid];
        if (!c->set || !tv_eq(v, &c->v)) {
            if (c->cap < n) { free(c->buf); c->buf = acq(n); c->cap = n; }
            c->v = *v;
            c->set = 0;
            vol(x, s->a, c->buf, v);
            c->set = 1;
        }
        memcpy(o, c->buf, sizeof(float) * (size_t)n);
        return;
    }
    case MCGK_BLEND_DENSITY: vol(x, s->a, o, v); return;
    case MCGK_INTERP: {
        const int cxz = s->i0, cy = s->i1;
        if ((v->dx == cxz || v->sx == 1) && (v->dy == cy || v->sy == 1) && (v->dz == cxz || v->sz == 1) &&
            floormod(v->x0, cxz) == 0 && floormod(v->y0, cy) == 0 && floormod(v->z0, cxz) == 0) {
            vol(x, s->a, o, v);
            return;
        }
        if (v->dx == 1 && v->dy == 1 && v->dz == 1) { interp_block(x, s, o, v); return; }
        const TV bv = { v->sx * v->dx, v->sy * v->dy, v->sz * v->dz, v->x0, v->y0, v->z0, 1, 1, 1 };
        float *bb = acq(tv_n(&bv));
        interp_block(x, s, bb, &bv);
        for (int iz = 0; iz < v->sz; iz++)
            for (int ix = 0; ix < v->sx; ix++)
                for (int iy = 0; iy < v->sy; iy++) o[tv_idx(v, ix, iy, iz)] = bb[tv_idx(&bv, ix * v->dx, iy * v->dy, iz * v->dz)];
        free(bb);
        return;
    }
    case MCGK_SLICE_X: case MCGK_SLICE_Y: case MCGK_SLICE_Z: case MCGK_SLICE_XZ: {
        TV iv = *v;
        if (s->k == MCGK_SLICE_X) { if (v->sx == 1 && v->x0 == s->i0) { vol(x, s->a, o, v); return; } iv.sx = 1; iv.x0 = s->i0; }
        if (s->k == MCGK_SLICE_Y) { if (v->sy == 1 && v->y0 == s->i0) { vol(x, s->a, o, v); return; } iv.sy = 1; iv.y0 = s->i0; }
        if (s->k == MCGK_SLICE_Z) { if (v->sz == 1 && v->z0 == s->i0) { vol(x, s->a, o, v); return; } iv.sz = 1; iv.z0 = s->i0; }
        if (s->k == MCGK_SLICE_XZ) {
            if (v->sx == 1 && v->sz == 1 && v->x0 == s->i0 && v->z0 == s->i1) { vol(x, s->a, o, v); return; }
            iv.sx = 1; iv.sz = 1; iv.x0 = s->i0; iv.z0 = s->i1;
        }
        float *ib = acq(tv_n(&iv));
        vol(x, s->a, ib, &iv);
        for (int iz = 0; iz < v->sz; iz++)
            for (int ix = 0; ix < v->sx; ix++)
                for (int iy = 0; iy < v->sy; iy++)
                    o[tv_idx(v, ix, iy, iz)] = ib[tv_idx(&iv, iv.sx == 1 ? 0 : ix, iv.sy == 1 ? 0 : iy, iv.sz == 1 ? 0 : iz)];
        free(ib);
        return;
    }
    default:
        x->bad = 1;
        fill(o, n, 0.0f);
        return;
    }
}

static void ctx_init(TCtx *x, const McgProg *P) {
    memset(x, 0, sizeof *x);
    x->P = P;
    x->ncell = P->ncache > 0 ? P->ncache : 0;
    x->cell = (TCell *)calloc((size_t)(x->ncell ? x->ncell : 1), sizeof(TCell));
}
static void ctx_free(TCtx *x) {
    for (int i = 0; i < x->ncell; i++) free(x->cell[i].buf);
    free(x->cell);
}

int or_terrain_volume(const McgProg *P, int node, const int vv[9], float *out) {
    TCtx x;
    ctx_init(&x, P);
    const TV v = { vv[0], vv[1], vv[2], vv[3], vv[4], vv[5], vv[6], vv[7], vv[8] };
    vol(&x, node, out, &v);
    const int bad = x.bad;
    ctx_free(&x);
    return bad ? -1 : 0;
}

/* Mth.getSeed (Java): (x * 3129871 as int) ^ z * 116129781L ^ y, then l * l * 42317861L + l * 11L, >> 16 */
static i64 get_seed(i32 x, i32 y, i32 z) {
    i64 l = (i64)(i32)((uint32_t)x * 3129871u) ^ ((i64)z * 116129781LL) ^ (i64)y;
    l = (i64)((u64)l * (u64)l * 42317861ULL + (u64)l * 11ULL);
    return l >> 16;
}
static u64 rotl(u64 v, int k) { return (v << k) | (v >> (64 - k)); }
typedef struct { u64 lo, hi; } Xo;
static u64 xo_next(Xo *r) {                       /* Xoroshiro128++ */
    const u64 s0 = r->lo;
    u64 s1 = r->hi;
    const u64 res = rotl(s0 + s1, 17) + s0;
    s1 ^= s0;
    r->lo = rotl(s0, 49) ^ s1 ^ (s1 << 21);
    r->hi = rotl(s1, 28);
    return res;
}
static float xo_float(Xo *r) { return (float)(xo_next(r) >> 40) * 5.9604645E-8f; }
static Xo xo_at(u64 lo, u64 hi, i32 x, i32 y, i32 z) {
    Xo r = { (u64)get_seed(x, y, z) ^ lo, hi };
    if ((r.lo | r.hi) == 0) { r.lo = 0x9E3779B97F4A7C15ULL; r.hi = 0x6A09E667F3BCC909ULL; }
    return r;
}
void or_terrain_ore_floats(uint64_t lo, uint64_t hi, int x, int y, int z, int k, float *out) {
    Xo r = xo_at(lo, hi, x, y, z);
    for (int i = 0; i < k; i++) out[i] = xo_float(&r);
}

int or_terrain_chunk(const McgTerrainWorld *T, int cx, int cz, int chunk_index, McgBeardFn beard, void *ud,
                     float *out_dens, uint16_t *out_veins, int *set, int *changed, int *cvol, float *cvals, int cap) {
    const McgProg *P = T->vol;
    TCtx x;
    ctx_init(&x, P);
    x.beard = beard; x.ud = ud; x.chunk = chunk_index;
    const int bx0 = cx * 16, bz0 = cz * 16;
    TV before[32];
    int had[32];
    const int nc = T->ncache_ids < 32 ? T->ncache_ids : 32;
    if (T->pre_root >= 0) {
        const TV pv = { T->pre_sx, T->pre_sy, T->pre_sz, bx0 + T->pre_xoff, T->pre_y0, bz0 + T->pre_zoff, T->pre_dx, T->pre_dy, T->pre_dz };
        float *tmp = acq(tv_n(&pv));
        vol(&x, P->root[T->pre_root], tmp, &pv);
        free(tmp);
    }
    for (int k = 0; k < nc; k++) {
        const int cid = T->cache_ids[k];
        had[k] = cid >= 0 && cid < x.ncell && x.cell[cid].set;
        if (had[k]) before[k] = x.cell[cid].v;
    }
    const TV v = { 16, T->nh, 16, bx0, T->nmin, bz0, 1, 1, 1 };
    vol(&x, P->root[T->root_density], out_dens, &v);
    for (int k = 0; k < nc; k++) {
        const int cid = T->cache_ids[k];
        set[k] = cid >= 0 && cid < x.ncell && x.cell[cid].set;
        changed[k] = set[k] && !(had[k] && tv_eq(&x.cell[cid].v, &before[k]));
        if (!set[k]) continue;
        const TV *cv = &x.cell[cid].v;
        const int c9[9] = { cv->sx, cv->sy, cv->sz, cv->x0, cv->y0, cv->z0, cv->dx, cv->dy, cv->dz };
        memcpy(cvol + 9 * k, c9, sizeof c9);
        const int nn = tv_n(cv) < cap ? tv_n(cv) : cap;
        memcpy(cvals + (size_t)k * (size_t)cap, x.cell[cid].buf, sizeof(float) * (size_t)nn);
    }
    if (out_veins && T->veins_on && T->nveins > 0) {
        const int nd = tv_n(&v);
        float *dens[MCG_MAX_VEINS], *rich[MCG_MAX_VEINS];
        for (int r = 0; r < T->nveins; r++) {
            dens[r] = acq(nd); rich[r] = acq(nd);
            vol(&x, P->root[T->vein[r].density_root], dens[r], &v);
            vol(&x, P->root[T->vein[r].richness_root], rich[r], &v);
        }
        for (int iz = 0; iz < 16; iz++)
            for (int ix = 0; ix < 16; ix++)
                for (int iy = 0; iy < T->nh; iy++) {
                    const int idx = tv_idx(&v, ix, iy, iz), bx = bx0 + ix, by = T->nmin + iy, bz = bz0 + iz;
                    uint16_t result = 0xFFFF;
                    for (int r = 0; r < T->nveins; r++) {
                        const float d = dens[r][idx];
                        if (d <= 0.0f) continue;
                        Xo rr = xo_at(T->ore_lo, T->ore_hi, bx, by, bz);
                        if (xo_float(&rr) > d) continue;
                        int outst;
                        int pass = xo_float(&rr) < rich[r][idx];
                        if (pass) pass = or_mcg_point_f(T->gap, T->gap->root[T->vein[r].gap_root], bx, by, bz) < 0.0f;
                        if (pass) outst = xo_float(&rr) < T->vein[r].raw_chance ? T->vein[r].raw : T->vein[r].ore;
                        else outst = T->vein[r].filler;
                        result = (uint16_t)outst;
                        break;
                    }
                    out_veins[(iy * 16 + iz) * 16 + ix] = result;
                }
        for (int r = 0; r < T->nveins; r++) { free(dens[r]); free(rich[r]); }
    }
    const int bad = x.bad;
    ctx_free(&x);
    return bad ? -1 : 0;
}
