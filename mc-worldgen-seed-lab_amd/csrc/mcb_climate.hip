// mcb_climate.hip - the engine-header entry points that return raw values instead of biomes:
//   k_climate_raw<MODE, DIM> : Climate.Sampler's six floats at block positions (engine/mc_climate.h:113-164;
//                              Nether: the two legacy NormalNoises of engine/cuda/mcg_host.c:63-74)
//   k_noise_values<MODE>     : one NormalNoise at arbitrary double positions (normal_get_d / normal_get_f,
//                              engine/mc_noise.h:286-324)
// Both take a lane per (seed, position), grid-stride, and read the seed's permutations and octave records
// from its slot (masked 256-byte tables); the gradient table and the offset spline are staged per
// workgroup.  The arithmetic is the grid kernels' (octave_term, normal2, column_finish), with a general
// three-coordinate octave for positions whose y is not the octave offset.
#include <algorithm>

#include "mcb_internal.h"
#include "mcb_device.h"

namespace mcb {
namespace {
using namespace dev;

// ImprovedNoise (double, engine/mc_noise.h:54-75) / GradientNoise (float, :90-110) at any (X, Y, Z):
// wrap (identity inside +-2^24), add the octave offsets, floor, fractions (cast to float for 26.3),
// 8 lattice hashes, then corners() in the reference's lerp nesting.
template <int MODE, class PA>
__device__ __forceinline__ typename R<MODE>::T oct_xyz(PA P, const typename GT<MODE>::T *G, const McbOct &o, double X,
                                                       double Y, double Z) {
    typedef typename R<MODE>::T T;
    const double x = wrapc(X) + o.xo, y = wrapc(Y) + o.yo, z = wrapc(Z) + o.zo;
    const double xf = floor(x), yf = floor(y), zf = floor(z);
    const int xi = (int)xf & 255, yi = (int)yf & 255, zi = (int)zf & 255;
    const T fx = (T)(x - xf), fy = (T)(y - yf), fz = (T)(z - zf);
    const int a = P[xi], b = P[xi + 1];
    const int aa = P[a + yi], ab = P[a + yi + 1], ba = P[b + yi], bb = P[b + yi + 1];
    const T u = MODE ? (T)fade_f((float)fx) : (T)fade_d((double)fx);
    const T v = MODE ? (T)fade_f((float)fy) : (T)fade_d((double)fy);
    const T w = MODE ? (T)fade_f((float)fz) : (T)fade_d((double)fz);
    return corners<MODE>(G, P[aa + zi], P[ba + zi], P[ab + zi], P[bb + zi], P[aa + zi + 1], P[ba + zi + 1], P[ab + zi + 1],
                         P[bb + zi + 1], fx, fy, fz, u, v, w);
}

// One NormalNoise at (X, Y, Z), its octaves at oct[0 .. 2*nlev): the reference's summation orders
// (26.3 interleaved first/second, or whole stacks for the legacy Nether noises; 26.1/26.2 two double
// stacks over wrapped coordinates, then (v1 + v2) * valueFactor).
template <int MODE>
__device__ typename R<MODE>::T normal3(SlotPerm perm, int base, const typename GT<MODE>::T *G, const McbOct *oct,
                                       const McbNoise &N, double X, double Y, double Z) {
    const int n = N.nlev;
    auto term = [&](int k, int st) -> typename R<MODE>::T {
        const McbOct &o = oct[st * n + k];
        const GP P = perm_at(perm, base + st * n + k);
        if constexpr (MODE == 1) {
            const double f = st ? N.freq[k] * INPUT_FACTOR : N.freq[k];
            return N.amp_f[k] * oct_xyz<1>(P, G, o, X * f, Y * f, Z * f);
        } else {
            const double Xs = st ? X * INPUT_FACTOR : X, Ys = st ? Y * INPUT_FACTOR : Y, Zs = st ? Z * INPUT_FACTOR : Z;
            const double f = N.freq[k];
            return N.amp_d[k] * oct_xyz<0>(P, G, o, wrapc(Xs * f), wrapc(Ys * f), wrapc(Zs * f)) * N.vf_d[k];
        }
    };
    if constexpr (MODE == 1) {
        float v = 0.0f;
        if (!N.legacy) {
            for (int k = 0; k < n; k++) {
                v += term(k, 0);
                v += term(k, 1);
            }
        } else {
            for (int k = 0; k < n; k++) v += term(k, 0);
            for (int k = 0; k < n; k++) v += term(k, 1);
        }
        return v;
    } else {
        double v1 = 0.0, v2 = 0.0;
        for (int k = 0; k < n; k++) v1 += term(k, 0);
        for (int k = 0; k < n; k++) v2 += term(k, 1);
        return (v1 + v2) * N.value_factor;
    }
}

// Raw climate of one block (engine/mc_climate.h:113-164): the shift noises at (x/4, 0, z/4) and
// (z/4, x/4, 0), the five climate noises at the shifted position with y = by * 0.0 (the octave offset),
// the offset from the spline, depth = y gradient + offset; floats in the order t, h, c, e, d, w.
template <int MODE, class PS>
__device__ __forceinline__ void ow_raw(const McbWorld *__restrict__ W, PS perm, const typename GT<MODE>::T *G,
                                       const SplLds *SP, const McbOct *oct, int bx, int by, int bz, float *out) {
    typedef typename R<MODE>::T T_;
    const McbNoise *N = W->noise;
    const double X = bx * 0.25, Z = bz * 0.25;
    const T_ a = normal2<MODE, false, 3, 0, false>(perm, G, oct, N[0], X, Z);
    const T_ b = normal2<MODE, true, 3, 0, false>(perm, G, oct, N[0], Z, X);
    double x, z;
    shift_coords<MODE>(a, b, X, Z, x, z);
    const T_ T = normal2<MODE, false, 2, 6, false>(perm, G, oct, N[1], x, z);
    const T_ H = normal2<MODE, false, 2, 10, false>(perm, G, oct, N[2], x, z);
    const T_ C = normal2<MODE, false, 9, 14, false>(perm, G, oct, N[3], x, z);
    const T_ E = normal2<MODE, false, 4, 32, false>(perm, G, oct, N[4], x, z);
    const T_ Wr = normal2<MODE, false, 3, 40, false>(perm, G, oct, N[5], x, z);
    Column c;
    column_finish<MODE>(W, SP, T, H, C, E, Wr, c);
    out[0] = (float)T; out[1] = (float)H; out[2] = (float)C; out[3] = (float)E;
    out[4] = ow_depth_raw<MODE>(W, c, by);
    out[5] = (float)Wr;
}

template <int MODE, int DIM>
__global__ __launch_bounds__(256) void k_climate_raw(const McbWorld *__restrict__ W, const McbSlot *__restrict__ slots, int ns,
                                                     const int32_t *__restrict__ xyz, int np, float *__restrict__ out) {
    __shared__ __attribute__((aligned(16))) typename GT<MODE>::T G[256];
    __shared__ SplLds spl[1];
    init_grad(G);
    if (DIM == MCB_OVERWORLD) stage_spline(spl, W->spline);
    __syncthreads();
    const int64_t n = (int64_t)ns * np;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const int s = (int)(i / np), p = (int)(i - (int64_t)s * np);
        const McbSlot *sl = slots + s;
        const int bx = xyz[3 * p], by = xyz[3 * p + 1], bz = xyz[3 * p + 2];
        float *o = out + 6 * i;
        if constexpr (DIM == MCB_OVERWORLD) {
            ow_raw<MODE>(W, SlotPerm{sl->perm}, G, spl, sl->oct, bx, by, bz, o);
        } else {
            // mcg_ref_points' Nether branch: (bx * 0.25, by * 0.0, bz * 0.25); y is the octave offset
            const double X = bx * 0.25, Z = bz * 0.25;
            o[0] = (float)normal2<MODE, false, 2, 0, true>(SlotPerm{sl->perm}, G, sl->oct, W->noise[0], X, Z);
            o[1] = (float)normal2<MODE, false, 2, 4, true>(SlotPerm{sl->perm}, G, sl->oct, W->noise[1], X, Z);
            o[2] = o[3] = o[4] = o[5] = 0.0f;
        }
    }
}

template <int MODE>
__global__ __launch_bounds__(256) void k_noise_values(const McbWorld *__restrict__ W, const McbSlot *__restrict__ slots, int ns,
                                                      int noise, const double *__restrict__ xyz, int np, double *__restrict__ out) {
    __shared__ __attribute__((aligned(16))) typename GT<MODE>::T G[256];
    init_grad(G);
    __syncthreads();
    const McbNoise &N = W->noise[noise];
    const int base = W->oct_base[noise];
    const int64_t n = (int64_t)ns * np;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const int s = (int)(i / np), p = (int)(i - (int64_t)s * np);
        const McbSlot *sl = slots + s;
        out[i] = (double)normal3<MODE>(SlotPerm{sl->perm}, base, G, sl->oct + base, N, xyz[3 * p], xyz[3 * p + 1],
                                       xyz[3 * p + 2]);
    }
}

int blocks_for(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 4096); }

}  // namespace

hipError_t launch_climate_raw(const DevWorld &w, const McbSlot *slots, int ns, const int32_t *xyz, int np, float *out,
                              hipStream_t st) {
    if (ns <= 0 || np <= 0) return hipSuccess;
    const dim3 g(blocks_for((int64_t)ns * np)), b(256);
    const bool f = w.host.fmt_new != 0;
    if (w.dim == MCB_OVERWORLD) {
        if (f) hipLaunchKernelGGL((k_climate_raw<1, MCB_OVERWORLD>), g, b, 0, st, w.d_world, slots, ns, xyz, np, out);
        else hipLaunchKernelGGL((k_climate_raw<0, MCB_OVERWORLD>), g, b, 0, st, w.d_world, slots, ns, xyz, np, out);
    } else if (w.dim == MCB_NETHER) {
        if (f) hipLaunchKernelGGL((k_climate_raw<1, MCB_NETHER>), g, b, 0, st, w.d_world, slots, ns, xyz, np, out);
        else hipLaunchKernelGGL((k_climate_raw<0, MCB_NETHER>), g, b, 0, st, w.d_world, slots, ns, xyz, np, out);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_noise_values(const DevWorld &w, const McbSlot *slots, int ns, int noise, const double *xyz, int np,
                               double *out, hipStream_t st) {
    if (ns <= 0 || np <= 0) return hipSuccess;
    if (w.dim == MCB_END || noise < 0 || noise >= w.host.n_noise) return hipErrorInvalidValue;
    const dim3 g(blocks_for((int64_t)ns * np)), b(256);
    if (w.host.fmt_new) hipLaunchKernelGGL((k_noise_values<1>), g, b, 0, st, w.d_world, slots, ns, noise, xyz, np, out);
    else hipLaunchKernelGGL((k_noise_values<0>), g, b, 0, st, w.d_world, slots, ns, noise, xyz, np, out);
    return hipGetLastError();
}

}  // namespace mcb
