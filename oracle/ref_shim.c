/* ref_shim.c - test infrastructure: exports of the reference's header-only CPU engine (engine/*.h,
 * included from REF_ROOT when `make -C oracle ref` builds _ref/libmcg_refshim.so).
 *
 * The reference's compiled driver (engine/cuda/mcg_host.c) exports only quantized targets and biomes.
 * Its raw climate floats (mc_climate_overworld_raw, engine/mc_climate.h:113-164) and NormalNoise values
 * (normal_get_f / normal_get_d, engine/mc_noise.h:286-324) are static inline functions; this file calls
 * them unchanged so tests can compare the oracle and the HIP engine with them bit for bit, including
 * where only rounding tells results apart (PerlinNoise.wrap past 2^24, engine/mc_noise.h:77,86-89).
 * Nothing here restates the reference; every value is computed by the reference's own code. */
#include "engine/mc_data.h"

typedef struct {
    McClimateSpec S;
    McNoiseSpec nether[2];
    int version;
} RsWorld;

static McVersionData g_V;    /* the reference loader fills one static (engine/cuda/mcg_host.c:11) */

/* the world of a version and Overworld preset ("overworld", "amplified", "large_biomes"); NULL on failure */
void *rs_open(int version, const char *preset) {
    if (mc_load_version(&g_V, version)) return NULL;
    McClimateSpec *S = mc_preset(&g_V, preset ? preset : "overworld");
    if (!S) return NULL;
    RsWorld *w = (RsWorld *)calloc(1, sizeof *w);
    w->S = *S;
    w->nether[0] = g_V.nether_temp;
    w->nether[1] = g_V.nether_veg;
    w->version = version;
    return w;
}

void rs_close(void *h) { free(h); }

/* mc_climate_overworld_raw for one seed at n block positions xyz[3i..3i+2] -> out[6i..6i+5] */
void rs_climate_raw(const void *h, int64_t seed, int n, const int32_t *xyz, float *out) {
    const RsWorld *w = (const RsWorld *)h;
    McClimate *cl = (McClimate *)malloc(sizeof *cl);
    mc_climate_init_overworld(cl, &w->S, seed);
    for (int i = 0; i < n; i++) mc_climate_overworld_raw(cl, &w->S, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], out + 6 * i);
    free(cl);
}

/* One NormalNoise of a seed at n double positions -> out[i] (float values widened to double for 26.3).
 * which: 0 offset, 1 temperature, 2 vegetation, 3 continentalness, 4 erosion, 5 ridge (the Overworld
 * climate noises, Xoroshiro-seeded as engine/mc_climate.h:97-106), 6 / 7 Nether temperature / vegetation
 * (legacy Random(seed) / Random(seed + 1), engine/cuda/mcg_host.c:44-46). */
int rs_noise(const void *h, int which, int64_t seed, int n, const double *xyz, double *out) {
    const RsWorld *w = (const RsWorld *)h;
    McNormal *nn = (McNormal *)malloc(sizeof *nn);
    const McNoiseSpec *s;
    if (which >= 0 && which < 6) {
        McClimate *cl = (McClimate *)malloc(sizeof *cl);
        mc_climate_init_overworld(cl, &w->S, seed);
        const McNormal *src[6] = {&cl->offset, &cl->temp, &cl->veg, &cl->cont, &cl->eros, &cl->ridge};
        const McNoiseSpec *sp[6] = {&w->S.sp_offset, &w->S.sp_temp, &w->S.sp_veg, &w->S.sp_cont, &w->S.sp_eros, &w->S.sp_ridge};
        *nn = *src[which];
        s = sp[which];
        free(cl);
    } else if (which == 6 || which == 7) {
        McLcg r = lcg_new(seed + (which == 7 ? 1 : 0));
        s = &w->nether[which - 6];
        normal_init_legacy(nn, s, &r);
    } else {
        free(nn);
        return -1;
    }
    for (int i = 0; i < n; i++) out[i] = normal_get(nn, s, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    free(nn);
    return 0;
}
