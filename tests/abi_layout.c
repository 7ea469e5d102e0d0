/* Prints the size and field offsets of every structure of the libmcgen GPU plugin ABI, and the values of
 * its enums, for the header named by ABI_HEADER.  tests/test_abi_layout.py compiles it once against
 * include/mcgpu.h and once against the reference's libmcgen/gpu/mcgen_gpu_abi.h and compares. */
#include <stddef.h>
#include <stdio.h>
#include ABI_HEADER

#define S(T) printf(#T " size %zu align %zu\n", sizeof(T), _Alignof(T))
#define F(T, f) printf(#T "." #f " %zu %zu\n", offsetof(T, f), sizeof(((T *)0)->f))
#define E(v) printf(#v " %d\n", (int)(v))

int main(void) {
    printf("MCGPU_ABI %d MCG_MAX_ROOTS %d\n", MCGPU_ABI, MCG_MAX_ROOTS);
    S(McgOct); F(McgOct, xo); F(McgOct, yo); F(McgOct, zo); F(McgOct, p);
    S(McgLayer); F(McgLayer, a); F(McgLayer, b); F(McgLayer, amp_d); F(McgLayer, amp_f); F(McgLayer, oct);
    F(McgLayer, kind); F(McgLayer, pad);
    S(McgNoise); F(McgNoise, l0); F(McgNoise, n0); F(McgNoise, l1); F(McgNoise, n1); F(McgNoise, value_factor);
    S(McgSpline); F(McgSpline, is_const); F(McgSpline, value); F(McgSpline, coord); F(McgSpline, n); F(McgSpline, loc0);
    F(McgSpline, val0);
    S(McgNode); F(McgNode, k); F(McgNode, a); F(McgNode, b); F(McgNode, c); F(McgNode, ns); F(McgNode, sp); F(McgNode, arr0);
    F(McgNode, narr); F(McgNode, thr0); F(McgNode, nthr); F(McgNode, i0); F(McgNode, i1); F(McgNode, i2); F(McgNode, i3);
    F(McgNode, f0); F(McgNode, f1); F(McgNode, f2); F(McgNode, f3); F(McgNode, d0); F(McgNode, d1); F(McgNode, cm);
    F(McgNode, cx); F(McgNode, cy); F(McgNode, cz); F(McgNode, oct); F(McgNode, cid);
    S(McgProg); F(McgProg, old); F(McgProg, nnodes); F(McgProg, nodes); F(McgProg, nnoise); F(McgProg, noise);
    F(McgProg, nlayers); F(McgProg, layers); F(McgProg, noct); F(McgProg, octs); F(McgProg, nsp); F(McgProg, sp);
    F(McgProg, nsplf); F(McgProg, splf); F(McgProg, nspc); F(McgProg, spc); F(McgProg, narr); F(McgProg, arr);
    F(McgProg, nthr); F(McgProg, thr); F(McgProg, nroots); F(McgProg, root); F(McgProg, ncache);
    S(McgTreeNode); F(McgTreeNode, lo); F(McgTreeNode, hi); F(McgTreeNode, first); F(McgTreeNode, count);
    F(McgTreeNode, biome); F(McgTreeNode, pad);
    S(McgBiomeWorld); F(McgBiomeWorld, source); F(McgBiomeWorld, fixed_biome); F(McgBiomeWorld, end_ids);
    F(McgBiomeWorld, newf); F(McgBiomeWorld, min_qy); F(McgBiomeWorld, qh); F(McgBiomeWorld, zoom_seed);
    F(McgBiomeWorld, prog); F(McgBiomeWorld, ntree); F(McgBiomeWorld, tree); F(McgBiomeWorld, tree_root);
#ifdef MCG_MAX_VEINS
    printf("MCG_MAX_VEINS %d\n", MCG_MAX_VEINS);
    S(McgVein); F(McgVein, density_root); F(McgVein, richness_root); F(McgVein, gap_root); F(McgVein, ore); F(McgVein, raw);
    F(McgVein, filler); F(McgVein, raw_chance);
    S(McgTerrainWorld); F(McgTerrainWorld, vol); F(McgTerrainWorld, gap); F(McgTerrainWorld, nmin); F(McgTerrainWorld, nh);
    F(McgTerrainWorld, root_density); F(McgTerrainWorld, nveins); F(McgTerrainWorld, vein); F(McgTerrainWorld, ore_lo);
    F(McgTerrainWorld, ore_hi); F(McgTerrainWorld, veins_on); F(McgTerrainWorld, cache_ids); F(McgTerrainWorld, ncache_ids);
    F(McgTerrainWorld, pre_root); F(McgTerrainWorld, pre_sx); F(McgTerrainWorld, pre_sy); F(McgTerrainWorld, pre_sz);
    F(McgTerrainWorld, pre_dx); F(McgTerrainWorld, pre_dy); F(McgTerrainWorld, pre_dz); F(McgTerrainWorld, pre_y0);
    F(McgTerrainWorld, pre_xoff); F(McgTerrainWorld, pre_zoff);
#endif
    E(MCGK_CONST); E(MCGK_NOISE); E(MCGK_NOISE_XZ); E(MCGK_NOISE_XYZ); E(MCGK_SHIFT_B); E(MCGK_END); E(MCGK_DIST);
    E(MCGK_GRAD_CLAMP); E(MCGK_GRAD_REPEAT); E(MCGK_GRAD_MIRROR); E(MCGK_CTX_ALPHA); E(MCGK_CTX_OFFSET); E(MCGK_CTX_BEARD);
    E(MCGK_ABS); E(MCGK_SQUARE); E(MCGK_CUBE); E(MCGK_SQRT); E(MCGK_LEAKY); E(MCGK_RECIP); E(MCGK_NEG); E(MCGK_SQUEEZE);
    E(MCGK_LOG); E(MCGK_SIGN); E(MCGK_ROUND_INT); E(MCGK_ROUND); E(MCGK_ADD); E(MCGK_CADD); E(MCGK_CSUB); E(MCGK_SUB);
    E(MCGK_MUL); E(MCGK_CMUL); E(MCGK_DIV); E(MCGK_CDIV); E(MCGK_MIN); E(MCGK_CMIN); E(MCGK_MAX); E(MCGK_CMAX);
    E(MCGK_POW_CB); E(MCGK_POW_CE); E(MCGK_POW); E(MCGK_SPLINE); E(MCGK_LERP); E(MCGK_LERP_CF); E(MCGK_LERP_CS);
    E(MCGK_CLAMP); E(MCGK_RANGE_C); E(MCGK_RANGE); E(MCGK_ISEL1); E(MCGK_ISEL); E(MCGK_CACHE); E(MCGK_BLEND_DENSITY);
    E(MCGK_INTERP); E(MCGK_SLICE_X); E(MCGK_SLICE_Y); E(MCGK_SLICE_Z); E(MCGK_SLICE_XZ); E(MCGK_FTS); E(MCGK__COUNT);
    E(MCGO_CONST); E(MCGO_NOISE); E(MCGO_SHIFTED_NOISE); E(MCGO_SHIFT_A); E(MCGO_SHIFT_B); E(MCGO_SHIFT);
    E(MCGO_END_ISLANDS); E(MCGO_Y_GRADIENT); E(MCGO_WEIRD_SCALED); E(MCGO_BLEND_ALPHA); E(MCGO_ZERO); E(MCGO_ABS);
    E(MCGO_SQUARE); E(MCGO_CUBE); E(MCGO_HALF_NEG); E(MCGO_QUARTER_NEG); E(MCGO_RECIP); E(MCGO_SQUEEZE); E(MCGO_CLAMP);
    E(MCGO_ADD); E(MCGO_MUL); E(MCGO_MIN); E(MCGO_MAX); E(MCGO_CADD); E(MCGO_CMUL); E(MCGO_RANGE); E(MCGO_ISEL);
    E(MCGO_SPLINE); E(MCGO_PASS); E(MCGO__COUNT);
    E(MCG_BS_MULTI); E(MCG_BS_END); E(MCG_BS_FIXED);
    return 0;
}
