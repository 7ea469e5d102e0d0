// mcg_router.cpp - recognition of the vanilla Overworld climate router in a plugin world.
//
// A libmcgen world hands the plugin its climate as a density-function program (include/mcgpu.h).  For the
// vanilla Overworld that program is a fixed graph: shift_x = NOISE(offset, xz 0.25, y 0) * 4 and shift_z =
// SHIFT_B(offset), five shifted climate noises, folded ridges, the offset spline over (continents,
// erosion, ridges, folded ridges) plus a constant, the blend (alpha 1 selects the offset itself) and depth
// = clamped y gradient + offset (26.3+), or the same graph in the 26.1/26.2 double lowering.  These are
// the operations of the engine's fused column code (mcb_device.h ow_noises / column_finish /
// ow_depth_raw; reference: engine/mc_climate.h:113-170), in the same order and precision, so when the
// graph matches node for node the plugin evaluates it with that code instead of interpreting the nodes.
// The graph is the one mcg_export.cpp writes (libmcgen's lowering, libmcgen/src/gpu_export_new.inc:
// 70-120, gpu_export_old.inc); caches and pass-through nodes are looked through.  Every constant is read
// from the world: the noise frequencies and amplitudes, the ridge, offset and gradient constants, the
// spline and the octave tables.  Anything else -- another graph, other level counts, a second-stack
// frequency that is not the first times 1.0181268882175227, slices, smeared layers -- returns false and
// the interpreter runs the program.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "mcg_plugin.h"

namespace mcb {
namespace plug {
namespace {

const double kInputFactor = 1.0181268882175227;
const int kLevels[6] = MCB_OW_LEVELS;          // offset, temperature, vegetation, continentalness, erosion, ridges
const int kBase[6] = {0, 6, 10, 14, 32, 40};  // first octave slot of each noise in the engine's seed layout

struct Match {
    const McgProg *p;
    bool neu;
    // the node a value comes from, through caches (26.3+) and pass-through nodes (26.1/26.2)
    int deref(int i) const {
        for (int guard = 0; i >= 0 && i < p->nnodes && guard < p->nnodes; guard++) {
            const McgNode &n = p->nodes[i];
            if (n.k != (neu ? MCGK_CACHE : MCGO_PASS)) return i;
            i = n.a;
        }
        return -1;
    }
    const McgNode *at(int i, int kind) const {
        i = deref(i);
        if (i < 0 || i >= p->nnodes) return nullptr;
        const McgNode *n = &p->nodes[i];
        return n->k == kind && n->cm == 0 ? n : nullptr;
    }
    bool same(int a, int b) const { return deref(a) >= 0 && deref(a) == deref(b); }
};

bool valid_oct(const McgProg *p, int o) { return o >= 0 && o < p->noct; }

// noise ns of the world as engine noise i (levels kLevels[i], octave slots from kBase[i])
bool take_noise(const McgProg *p, bool neu, int ns, int i, McbWorld &w, McbSlot &slot) {
    if (ns < 0 || ns >= p->nnoise) return false;
    const McgNoise &mn = p->noise[ns];
    const int nl = kLevels[i];
    McbNoise &N = w.noise[i];
    N.nlev = nl;
    auto put = [&](int o, const McgOct &m) {
        memcpy(slot.perm[o], m.p, 256);
        slot.oct[o].xo = m.xo;
        slot.oct[o].yo = m.yo;
        slot.oct[o].zo = m.zo;
    };
    if (neu) {
        // one stack, levels interleaved: layer 2k = level k, layer 2k+1 = level k at frequency * 1.0181...
        if (mn.n0 != 2 * nl || mn.l0 < 0 || mn.l0 + mn.n0 > p->nlayers) return false;
        for (int k = 0; k < nl; k++) {
            const McgLayer &a = p->layers[mn.l0 + 2 * k], &b = p->layers[mn.l0 + 2 * k + 1];
            if (a.kind != 0 || b.kind != 0 || !valid_oct(p, a.oct) || !valid_oct(p, b.oct)) return false;
            if (!(b.a == a.a * kInputFactor) || !(a.amp_f == b.amp_f)) return false;
            N.freq[k] = a.a;
            N.amp_f[k] = a.amp_f;
            put(kBase[i] + k, p->octs[a.oct]);
            put(kBase[i] + nl + k, p->octs[b.oct]);
        }
    } else {
        // two PerlinNoises of nl levels each, the second at the coordinates * 1.0181...
        if (mn.n0 != nl || mn.n1 != nl || mn.l0 < 0 || mn.l1 < 0 || mn.l0 + nl > p->nlayers || mn.l1 + nl > p->nlayers)
            return false;
        for (int k = 0; k < nl; k++) {
            const McgLayer &a = p->layers[mn.l0 + k], &b = p->layers[mn.l1 + k];
            if (!valid_oct(p, a.oct) || !valid_oct(p, b.oct)) return false;
            if (!(a.a == b.a) || !(a.b == b.b) || !(a.amp_d == b.amp_d)) return false;
            N.freq[k] = a.a;
            N.vf_d[k] = a.b;
            N.amp_d[k] = a.amp_d;
            put(kBase[i] + k, p->octs[a.oct]);
            put(kBase[i] + nl + k, p->octs[b.oct]);
        }
        N.value_factor = mn.value_factor;
    }
    return true;
}

// spline node si of the world into the engine's flattened form (node 0 = the root); coord[j] maps the
// spline node's coordinate j to the engine's (0 continents, 1 erosion, 2 ridges, 3 folded ridges)
int take_spline(const McgProg *p, int si, const int coord[], int ncoord, McbSpline &S, int level, int *depth) {
    if (si < 0 || si >= p->nsp || S.n_nodes >= MCB_SPL_NODES || level >= MCB_SPL_MAXD) return -1;
    const McgSpline &m = p->sp[si];
    const int id = S.n_nodes++;
    if (m.is_const) {
        S.kind[id] = 0;
        S.value[id] = m.value;
        S.hdr[id] = 0;
        return id;
    }
    if (m.coord < 0 || m.coord >= ncoord || coord[m.coord] < 0 || m.n < 1 || m.n > 255) return -1;
    if (m.loc0 < 0 || m.loc0 + 2 * m.n > p->nsplf || m.val0 < 0 || m.val0 + m.n > p->nspc) return -1;
    const int first = S.n_pts;
    if (first + m.n > MCB_SPL_PTS) return -1;
    S.n_pts += m.n;
    S.kind[id] = 1;
    S.coord[id] = coord[m.coord];
    S.count[id] = m.n;
    S.first[id] = first;
    S.hdr[id] = 1u << 31 | (uint32_t)S.coord[id] << 24 | (uint32_t)m.n << 16 | (uint32_t)first;
    *depth = std::max(*depth, level + 1);
    for (int i = 0; i < m.n; i++) {
        S.loc[first + i] = p->splf[m.loc0 + i];
        S.der[first + i] = p->splf[m.loc0 + m.n + i];
        const int c = take_spline(p, p->spc[m.val0 + i], coord, ncoord, S, level + 1, depth);
        if (c < 0) return -1;
        S.child[first + i] = c;
    }
    return id;
}

}  // namespace

bool vanilla_router(const McgBiomeWorld *W, McbWorld &w, McbSlot &slot) {
    memset(&w, 0, sizeof w);
    memset(&slot, 0, sizeof slot);
    if (!W || W->source != MCG_BS_MULTI || !W->prog || W->prog->nroots < 6) return false;
    const McgProg *p = W->prog;
    const bool neu = !p->old;
    if (neu != (W->newf != 0)) return false;
    const Match M{p, neu};
    // roots: 0 temperature, 1 vegetation, 2 continents, 3 erosion, 4 depth, 5 ridges
    const int field_root[5] = {0, 1, 2, 3, 5};
    const McgNode *f[5];
    for (int i = 0; i < 5; i++) {
        f[i] = M.at(p->root[field_root[i]], neu ? MCGK_NOISE_XZ : MCGO_SHIFTED_NOISE);
        if (!f[i] || !(f[i]->d0 == 0.25) || !(f[i]->d1 == 0.0)) return false;
        if (!M.same(f[i]->a, f[0]->a) || !M.same(f[i]->c, f[0]->c)) return false;
        if (!neu) {                                               // y shift: the constant 0
            const McgNode *z = M.at(f[i]->b, MCGO_CONST);
            if (!z || !(z->d0 == 0.0) || std::signbit(z->d0)) return false;
        }
    }
    // shift_x / shift_z of the offset noise
    int ns0;
    if (neu) {
        const McgNode *sx = M.at(f[0]->a, MCGK_CMUL), *sz = M.at(f[0]->c, MCGK_SHIFT_B);
        if (!sx || !sz || !(sx->f0 == 4.0f) || !(sz->d0 == 0.25)) return false;
        const McgNode *a = M.at(sx->a, MCGK_NOISE);
        if (!a || !(a->d0 == 0.25) || !(a->d1 == 0.0) || a->ns != sz->ns) return false;
        ns0 = a->ns;
    } else {
        const McgNode *sx = M.at(f[0]->a, MCGO_SHIFT_A), *sz = M.at(f[0]->c, MCGO_SHIFT_B);
        if (!sx || !sz || !(sx->d0 == 0.25) || !(sz->d0 == 0.25) || sx->ns != sz->ns) return false;
        ns0 = sx->ns;
    }
    w.version = neu ? MCB_V26_3 : MCB_V26_1;
    w.dim = MCB_OVERWORLD;
    w.fmt_new = neu;
    w.n_noise = 6;
    w.n_oct = MCB_OW_OCT;
    for (int i = 0; i < 6; i++) w.oct_base[i] = kBase[i];
    if (!take_noise(p, neu, ns0, 0, w, slot)) return false;
    for (int i = 0; i < 5; i++)
        if (!take_noise(p, neu, f[i]->ns, 1 + i, w, slot)) return false;
    const int C = M.deref(p->root[2]), E = M.deref(p->root[3]), Wr = M.deref(p->root[5]);
    // folded ridges (|(|W| + r0)| + r1) * r2
    const int k_abs = neu ? MCGK_ABS : MCGO_ABS, k_cadd = neu ? MCGK_CADD : MCGO_CADD, k_cmul = neu ? MCGK_CMUL : MCGO_CMUL;
    auto cval = [&](const McgNode *n) { return neu ? (double)n->f0 : n->d0; };
    // depth = gradient + offset (either operand order)
    const McgNode *dep = M.at(p->root[4], neu ? MCGK_ADD : MCGO_ADD);
    if (!dep) return false;
    const McgNode *grad = M.at(dep->a, neu ? MCGK_GRAD_CLAMP : MCGO_Y_GRADIENT);
    int offv = dep->b;
    if (!grad) {
        grad = M.at(dep->b, neu ? MCGK_GRAD_CLAMP : MCGO_Y_GRADIENT);
        offv = dep->a;
    }
    if (!grad) return false;
    const McgNode *spl;
    if (neu) {
        if (grad->i0 != 1 || grad->i2 > grad->i3 || (grad->i1 != grad->i2 && grad->i1 != grad->i3)) return false;
        w.gy0 = grad->i1;
        w.gy1 = grad->i1 == grad->i2 ? grad->i3 : grad->i2;
        w.gv0_f = grad->f0;
        w.grad_factor_f = grad->f1;
        // the blend: lerp(alpha 1, blend offset, offset) is the offset itself
        const McgNode *le = M.at(offv, MCGK_LERP);
        if (!le) return false;
        const McgNode *al = M.at(le->a, MCGK_CTX_ALPHA);
        if (!al || !(al->f0 == 1.0f) || !M.at(le->b, MCGK_CTX_OFFSET)) return false;
        const McgNode *off = M.at(le->c, MCGK_CADD);
        if (!off) return false;
        w.offc_f = off->f0;
        spl = M.at(off->a, MCGK_SPLINE);
    } else {
        w.gy0 = grad->i0;
        w.gy1 = grad->i1;
        w.gv0_d = grad->d0;
        w.gv1_d = grad->d1;
        // ((spline + offset constant) * 1) + 0
        const McgNode *o0 = M.at(offv, MCGO_CADD);
        if (!o0 || !(o0->d0 == 0.0) || std::signbit(o0->d0)) return false;
        const McgNode *o1 = M.at(o0->a, MCGO_CMUL);
        if (!o1 || !(o1->d0 == 1.0)) return false;
        const McgNode *o2 = M.at(o1->a, MCGO_CADD);
        if (!o2) return false;
        w.offc_d = o2->d0;
        spl = M.at(o2->a, MCGO_SPLINE);
    }
    if (!spl || spl->narr < 1 || spl->arr0 < 0 || spl->arr0 + spl->narr > p->narr) return false;
    // the spline's fourth coordinate: folded ridges over the ridge root
    int coord[64];
    if (spl->narr > 64) return false;
    bool have_rf = false;
    for (int j = 0; j < spl->narr; j++) {
        const int a = M.deref(p->arr[spl->arr0 + j]);
        coord[j] = a == C ? 0 : (a == E ? 1 : (a == Wr ? 2 : -1));
        if (coord[j] >= 0) continue;
        const McgNode *m2 = M.at(a, k_cmul);
        const McgNode *a1 = m2 ? M.at(m2->a, k_cadd) : nullptr;
        const McgNode *b1 = a1 ? M.at(a1->a, k_abs) : nullptr;
        const McgNode *a0 = b1 ? M.at(b1->a, k_cadd) : nullptr;
        const McgNode *b0 = a0 ? M.at(a0->a, k_abs) : nullptr;
        if (!b0 || M.deref(b0->a) != Wr) continue;
        const double r[3] = {cval(a0), cval(a1), cval(m2)};
        if (have_rf) {                                            // the engine has one folded-ridge value
            if (memcmp(r, w.rf_d, sizeof r) != 0) return false;
        }
        for (int q = 0; q < 3; q++) {
            w.rf_d[q] = r[q];
            w.rf_f[q] = (float)r[q];
        }
        have_rf = true;
        coord[j] = 3;
    }
    int depth = 0;
    if (take_spline(p, spl->sp, coord, spl->narr, w.spline, 0, &depth) != 0) return false;
    w.spline.depth = depth;
    return true;
}

}  // namespace plug
}  // namespace mcb
