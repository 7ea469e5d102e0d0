/* ref_terrain.c - test infrastructure: libmcgen's own 26.3+ density-function pipeline, compiled from its
 * sources where they lie (libmcgen/src/df_new.c with gpu_export_new.inc, df_parse.c, json.c, noise.c,
 * util.c; `make -C oracle ref` builds _ref/libmcgen_ref.so), driven over worlds described in the game's
 * datapack JSON by the tests: a world seed, noise parameters by name and density functions by id.
 *
 * What it gives the tests, all computed by the reference's code:
 *   - the compiled sampler tree of a density function (nc_get, df_new.c:598-607) and its GPU program
 *     (gpu_export_new, gpu_export_new.inc:148-160) - the McgProg a libmcgen host hands mcgpu_terrain_new;
 *   - sampleVolume over a volume (s_volume, df_new.c:917-1150) with the sampler context's caches, as the
 *     CPU terrain path runs it per chunk, and the cache cells left behind (gpu_sctx_get_cell);
 *   - sampleValue at points (s_value, df_new.c:740-862);
 *   - the ore PositionalRandomFactory and its per-block floats (noise.c:16-30, world.c:182-183).
 * The noise environment follows libmcgen's world set-up (world.c:40-58, 83-106, 170-184): noises from
 * the terrain positional factory of Xoroshiro(seed) by name, the End-islands simplex from
 * Legacy(seed) after 17292 draws.  Nothing here restates the reference. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "df_new.h"
#include "gpu_bridge.h"
#include "json.h"

#define RT_MAX_ROOTS 32

typedef struct {
    char *name;
    NoiseParams P;
    NStack ns;
    int ready;
} RtNoise;

typedef struct {
    char *id;
    Df *df;
} RtFunc;

typedef struct {
    i64 seed;
    PosRnd pos_terrain, ore_pos;
    JsDoc *noises_doc, *funcs_doc;
    RtNoise *noise; int nnoise;
    RtFunc *func; int nfunc;
    Df *owned[RT_MAX_ROOTS]; int nowned;
    GNoise end_simplex; int has_end;
    BlendFbm blended[4]; const Df *blended_key[4]; int nblended;
    NComp *nc;
    const S *root[RT_MAX_ROOTS]; int nroots;
    McgProg prog[2]; int has_prog[2];
} RtWorld;

static const NStack *env_noise(void *ud, const char *name, char *err, size_t errlen) {
    RtWorld *w = (RtWorld *)ud;
    for (int i = 0; i < w->nnoise; i++) {
        RtNoise *n = &w->noise[i];
        if (strcmp(n->name, name)) continue;
        if (!n->ready) {
            Rnd r = pos_from_hash(&w->pos_terrain, name);
            nn_new_create(&n->ns, &n->P, &r);
            n->ready = 1;
        }
        return &n->ns;
    }
    snprintf(err, errlen, "no noise %s", name);
    return NULL;
}
static const BlendFbm *env_blended(void *ud, const Df *f) {
    RtWorld *w = (RtWorld *)ud;
    for (int i = 0; i < w->nblended; i++) if (df_equal(w->blended_key[i], f)) return &w->blended[i];
    if (w->nblended >= 4) return &w->blended[0];
    Rnd r = pos_from_hash(&w->pos_terrain, "minecraft:terrain");
    blended_new_create(&w->blended[w->nblended], &r, f->n1.d, f->n4.d, f->n3.d);
    w->blended_key[w->nblended] = f;
    return &w->blended[w->nblended++];
}
static const GNoise *env_end(void *ud) {
    RtWorld *w = (RtWorld *)ud;
    if (!w->has_end) {
        Rnd r = rnd_legacy_seed(w->seed);
        rnd_consume(&r, 17292);
        gn_init(&w->end_simplex, &r, 0.0);
        w->has_end = 1;
    }
    return &w->end_simplex;
}
static const Df *env_ref(void *ud, const char *id) {
    RtWorld *w = (RtWorld *)ud;
    for (int i = 0; i < w->nfunc; i++) if (!strcmp(w->func[i].id, id)) return w->func[i].df;
    return NULL;
}
static Ival env_range(void *ud, const char *name) {
    RtWorld *w = (RtWorld *)ud;
    for (int i = 0; i < w->nnoise; i++) if (!strcmp(w->noise[i].name, name)) return nn_new_range(&w->noise[i].P);
    return iv_inf();
}

static void prog_release(McgProg *p) {
    free(p->nodes); free(p->noise); free(p->layers); free(p->octs); free(p->sp); free(p->splf); free(p->spc);
    free(p->arr); free(p->thr);
    memset(p, 0, sizeof *p);
}

void rt_close(void *h) {
    RtWorld *w = (RtWorld *)h;
    if (!w) return;
    for (int k = 0; k < 2; k++) if (w->has_prog[k]) prog_release(&w->prog[k]);
    if (w->nc) nc_free(w->nc);
    for (int i = 0; i < w->nnoise; i++) { if (w->noise[i].ready) ns_free(&w->noise[i].ns); free(w->noise[i].name); }
    for (int i = 0; i < w->nblended; i++) { ns_free(&w->blended[i].min_lim); ns_free(&w->blended[i].max_lim); ns_free(&w->blended[i].main); }
    for (int i = 0; i < w->nfunc; i++) { df_free(w->func[i].df); free(w->func[i].id); }
    for (int i = 0; i < w->nowned; i++) df_free(w->owned[i]);
    free(w->noise); free(w->func);
    if (w->noises_doc) js_free(w->noises_doc);
    if (w->funcs_doc) js_free(w->funcs_doc);
    free(w);
}

/* A world: seed, noises {"minecraft:id": {noise parameters}}, density functions {"minecraft:id": df}.
 * NULL on failure with the reason in err. */
void *rt_open(int64_t seed, const char *noises_json, const char *funcs_json, char *err, size_t errlen) {
    RtWorld *w = (RtWorld *)calloc(1, sizeof *w);
    w->seed = seed;
    Rnd rt = rnd_xoro_seed(seed);
    w->pos_terrain = rnd_fork_positional(&rt);
    Rnd o = pos_from_hash(&w->pos_terrain, "minecraft:ore");
    w->ore_pos = rnd_fork_positional(&o);
    w->noises_doc = js_parse(noises_json, strlen(noises_json), err, errlen);
    w->funcs_doc = w->noises_doc ? js_parse(funcs_json, strlen(funcs_json), err, errlen) : NULL;
    if (!w->noises_doc || !w->funcs_doc) { rt_close(w); return NULL; }
    const Js *nj = js_root(w->noises_doc), *fj = js_root(w->funcs_doc);
    if (!js_is_obj(nj) || !js_is_obj(fj)) { snprintf(err, errlen, "noises and functions must be objects"); rt_close(w); return NULL; }
    w->noise = (RtNoise *)calloc((size_t)(nj->n ? nj->n : 1), sizeof(RtNoise));
    for (int i = 0; i < nj->n; i++) {
        RtNoise *n = &w->noise[w->nnoise++];
        n->name = df_full_id(nj->keys[i]);
        if (noise_params_parse(nj->items[i], 1, &n->P, err, errlen)) { rt_close(w); return NULL; }
    }
    w->func = (RtFunc *)calloc((size_t)(fj->n ? fj->n : 1), sizeof(RtFunc));
    for (int i = 0; i < fj->n; i++) {
        RtFunc *f = &w->func[w->nfunc++];
        f->id = df_full_id(fj->keys[i]);
        f->df = df_parse(fj->items[i], 1, err, errlen);
        if (!f->df) { rt_close(w); return NULL; }
    }
    NEnv env = { w, env_noise, env_blended, env_end, env_ref, env_range, 0, NULL };
    w->nc = nc_new(&env);
    return w;
}

/* compile a density function (JSON text: an object, a number or an id string) as the next root; its index,
 * or -1 with the reason in err */
int rt_compile(void *h, const char *df_json, char *err, size_t errlen) {
    RtWorld *w = (RtWorld *)h;
    if (w->nroots >= RT_MAX_ROOTS || w->nowned >= RT_MAX_ROOTS) { snprintf(err, errlen, "too many roots"); return -1; }
    JsDoc *d = js_parse(df_json, strlen(df_json), err, errlen);
    if (!d) return -1;
    Df *f = df_parse(js_root(d), 1, err, errlen);
    js_free(d);
    if (!f) return -1;
    w->owned[w->nowned++] = f;
    const S *s = nc_get(w->nc, f, err, errlen);
    if (!s) return -1;
    w->root[w->nroots] = s;
    return w->nroots++;
}

/* the GPU program of roots[0..n) (slot 0: volume program, bake_ctx 0; slot 1: point program, bake_ctx 1) */
const McgProg *rt_export(void *h, int slot, const int *roots, int n) {
    RtWorld *w = (RtWorld *)h;
    if (slot < 0 || slot > 1 || n < 0 || n > MCG_MAX_ROOTS) return NULL;
    if (w->has_prog[slot]) prog_release(&w->prog[slot]);
    const S *r[MCG_MAX_ROOTS];
    for (int i = 0; i < n; i++) r[i] = (roots[i] >= 0 && roots[i] < w->nroots) ? w->root[roots[i]] : NULL;
    if (gpu_export_new(r, n, slot, &w->prog[slot])) return NULL;
    w->has_prog[slot] = 1;
    return &w->prog[slot];
}

/* cache ids of a root's sampler tree (gpu_collect_cache_ids), appended to ids[*n..cap) */
void rt_cache_ids(void *h, int root, int *ids, int cap, int *n) {
    RtWorld *w = (RtWorld *)h;
    if (root >= 0 && root < w->nroots) gpu_collect_cache_ids(w->root[root], ids, cap, n);
}

/* One chunk as the CPU terrain path samples it: a fresh sampler context (caches on), optionally the
 * pre-density volume of root `pre` over pre_vol first, then root `dens` over vol into out.  For each of
 * the ncid cache ids: set[k] = 1 and cvol[9k..] = the cell's volume when the cell holds a buffer after
 * the density, changed[k] = 1 when that differs from the state after the pre volume; the values go to
 * cvals + k * cap (at most cap each).  Volumes are {sx, sy, sz, x0, y0, z0, dx, dy, dz}. */
void rt_chunk(void *h, int pre, const int *pre_vol, int dens, const int *vol, float *out, int ncid, const int *cids,
              int *set, int *changed, int *cvol, float *cvals, int cap) {
    RtWorld *w = (RtWorld *)h;
    SCtx *x = sctx_new(w->nc, 1);
    Vol before[64]; int had[64];
    if (ncid > 64) ncid = 64;
    if (pre >= 0) {
        const Vol pv = { pre_vol[0], pre_vol[1], pre_vol[2], pre_vol[3], pre_vol[4], pre_vol[5], pre_vol[6], pre_vol[7], pre_vol[8] };
        float *tmp = (float *)malloc(sizeof(float) * (size_t)vol_size(&pv));
        s_volume(x, w->root[pre], tmp, &pv);
        free(tmp);
    }
    for (int k = 0; k < ncid; k++) {
        const float *vals;
        had[k] = gpu_sctx_get_cell(x, cids[k], &before[k], &vals);
    }
    const Vol v = { vol[0], vol[1], vol[2], vol[3], vol[4], vol[5], vol[6], vol[7], vol[8] };
    s_volume(x, w->root[dens], out, &v);
    for (int k = 0; k < ncid; k++) {
        Vol cv; const float *vals = NULL;
        set[k] = gpu_sctx_get_cell(x, cids[k], &cv, &vals);
        changed[k] = set[k] && !(had[k] && vol_eq(&cv, &before[k]));
        if (!set[k]) continue;
        const int cvv[9] = { cv.sx, cv.sy, cv.sz, cv.x0, cv.y0, cv.z0, cv.dx, cv.dy, cv.dz };
        memcpy(cvol + 9 * k, cvv, sizeof cvv);
        const int nn = vol_size(&cv) < cap ? vol_size(&cv) : cap;
        memcpy(cvals + (size_t)k * (size_t)cap, vals, sizeof(float) * (size_t)nn);
    }
    sctx_free(x);
}

/* sampleVolume of a root over one volume in a fresh context; caches on or off */
void rt_volume(void *h, int root, const int *vol, int caches, float *out) {
    RtWorld *w = (RtWorld *)h;
    SCtx *x = sctx_new(w->nc, caches);
    const Vol v = { vol[0], vol[1], vol[2], vol[3], vol[4], vol[5], vol[6], vol[7], vol[8] };
    s_volume(x, w->root[root], out, &v);
    sctx_free(x);
}

/* sampleValue of a root at n block points xyz[3i..] (no caches) */
void rt_values(void *h, int root, int n, const int *xyz, float *out) {
    RtWorld *w = (RtWorld *)h;
    SCtx *x = sctx_new(w->nc, 0);
    for (int i = 0; i < n; i++) out[i] = s_value(x, w->root[root], xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    sctx_free(x);
}

/* the ore PositionalRandomFactory (lo, hi) and the first k floats of its random at each block point */
void rt_ore(void *h, uint64_t *lo, uint64_t *hi) {
    RtWorld *w = (RtWorld *)h;
    *lo = w->ore_pos.lo;
    *hi = w->ore_pos.hi;
}
void rt_ore_floats(void *h, int n, const int *xyz, int k, float *out) {
    RtWorld *w = (RtWorld *)h;
    for (int i = 0; i < n; i++) {
        Rnd r = pos_at(&w->ore_pos, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
        for (int j = 0; j < k; j++) out[(size_t)i * (size_t)k + (size_t)j] = rnd_next_float(&r);
    }
}
