// mcg_terrain.hip - kernels of the terrain volume executor (mcgpu_terrain_*; plan in mcg_terrain.cpp).
//
// k_tpass runs one pass: lane = (chunk, element) of the pass's volume, flattened so that small volumes
// (sliced 2D functions, cell corners) still fill whole waves.  Every node kind follows sampleVolume
// (libmcgen/src/df_new.c:917-1150) with its operation order: the NoiseStack volume order (noise.c:141-190,
// 359-366: the x and z gradient terms are summed first and the y term added, coordinates are
// block * (scale * frequency), the amplitude is amp * layer amplitude), the transposed volume of shift_b,
// the volume path's element-wise arithmetic (plain comparisons in min / max, no zero short-cut in mul /
// div), point NoiseStack values for shifted noises, the cell-step interpolation with its running y sum
// (df_new.c:867-915) and the slices.  The reference's per-node kernels are
// libmcgen/gpu/mcgpu_terrain.cu:38-328.  Builds use -ffp-contract=off; the fused multiply-adds written
// below multiply by a gradient component in {-1, 0, 1}, an exact product, so they round like the separate
// product and sum.
//
// k_tveins applies the ore-vein rules per block (mcgpu_terrain.cu:303-324) and k_tgather copies the
// cache cells the CPU needs into one staging buffer.
#include "mcg_plugin_dev.h"
#include "mcg_terrain.h"

namespace mcb {
namespace plug {

// Records every lane of a wave reads at the same address (the instruction, a noise's layer list, an
// octave's offsets) are read through the constant address space, i.e. with scalar loads, once per wave.
template <typename T>
__device__ __forceinline__ T cload(const T *p) {
#if __HIP_DEVICE_COMPILE__
    return *(const __attribute__((address_space(4))) T *)p;
#else
    return *p;                                   // host pass: never executed
#endif
}

__device__ __forceinline__ int tvx0(const TVol &v, int cx0) { return v.xrel ? cx0 + v.xoff : v.xoff; }
__device__ __forceinline__ int tvz0(const TVol &v, int cz0) { return v.zrel ? cz0 + v.zoff : v.zoff; }

// one octave of NoiseStack.addToVolume at one element (noise.c:141-190)
__device__ __forceinline__ float vol_octave(const McgOct *__restrict__ o, const GradEnt *G, int smeared, double fs, int bx,
                                            int by, int bz, double sxz, double sy) {
    const uint8_t *P = o->p;
    const double xo = cload(&o->xo), yo = cload(&o->yo), zo = cload(&o->zo);
    const double z = wrap_new(bz * sxz) + zo;
    const int32_t fz = jfloor(z);
    const float rz = (float)(z - fz), az = smooth_f(rz);
    const double x = wrap_new(bx * sxz) + xo;
    const int32_t fx = jfloor(x);
    const float rx = (float)(x - fx), ax = smooth_f(rx);
    const double oy = by * sy;
    const double y = wrap_new(oy) + yo;
    const int32_t fy = jfloor(y);
    const double ryd = y - fy;
    const float ay = smooth_f((float)ryd);
    float ry = (float)ryd;
    if (smeared) {
        const double lim = (oy >= 0.0 && oy < ryd) ? oy : ryd;
        ry = (float)(ryd - (double)jfloor(lim / fs + (double)1.0E-7f) * fs);
    }
    const int x0 = P[fx & 255], x1 = P[(fx + 1) & 255];
    const int xy00 = P[(x0 + fy) & 255], xy01 = P[(x0 + fy + 1) & 255];
    const int xy10 = P[(x1 + fy) & 255], xy11 = P[(x1 + fy + 1) & 255];
    const float gx = rx - 1.0f, gy = ry - 1.0f, gz = rz - 1.0f;
    // (g.x * x + g.z * z) + g.y * y
    auto vdot = [](GradEnt g, float xx, float yy, float zz) {
        const float cx = __uint_as_float((uint32_t)g);
        const dev::half2v yz = __builtin_bit_cast(dev::half2v, (uint32_t)(g >> 32));
        return fmaf((float)yz.x, yy, fmaf((float)yz.y, zz, cx * xx));
    };
    const float c000 = vdot(G[P[(xy00 + fz) & 255]], rx, ry, rz), c100 = vdot(G[P[(xy10 + fz) & 255]], gx, ry, rz);
    const float c010 = vdot(G[P[(xy01 + fz) & 255]], rx, gy, rz), c110 = vdot(G[P[(xy11 + fz) & 255]], gx, gy, rz);
    const float c001 = vdot(G[P[(xy00 + fz + 1) & 255]], rx, ry, gz), c101 = vdot(G[P[(xy10 + fz + 1) & 255]], gx, ry, gz);
    const float c011 = vdot(G[P[(xy01 + fz + 1) & 255]], rx, gy, gz), c111 = vdot(G[P[(xy11 + fz + 1) & 255]], gx, gy, gz);
    return jlerpf(az, jlerpf(ay, jlerpf(ax, c000, c100), jlerpf(ax, c010, c110)),
                  jlerpf(ay, jlerpf(ax, c001, c101), jlerpf(ax, c011, c111)));
}
// NoiseStack.addToVolume into a zero buffer element (noise.c:359-366)
__device__ float vol_stack(const TProg &P, const GradEnt *G, int nsi, int bx, int by, int bz, double sxz, double sy, float amp) {
    const McgNoise mn = cload(P.noise + nsi);
    float buf = 0.0f;
#pragma unroll 1
    for (int l = 0; l < mn.n0; l++) {
        const McgLayer L = cload(P.layers + mn.l0 + l);
        const double f = L.a;
        buf += (amp * L.amp_f) * vol_octave(P.octs + L.oct, G, L.kind, L.b, bx, by, bz, sxz * f, sy * f);
    }
    return buf;
}
// NoiseStack.get (noise.c:339-347), the point path of shifted noises
__device__ float pt_stack(const TProg &P, const GradEnt *G, int nsi, double x, double y, double z) {
    const McgNoise mn = cload(P.noise + nsi);
    float v = 0.0f;
#pragma unroll 1
    for (int i = 0; i < mn.n0; i++) {
        const McgLayer L = cload(P.layers + mn.l0 + i);
        const McgOct *o = P.octs + L.oct;
        const double f = L.a;
        const Off3 off{cload(&o->xo), cload(&o->yo), cload(&o->zo)};
        v += L.amp_f * octave_f<false>(o->p, G, &off, L.kind, L.b, x * f, y * f, z * f);
    }
    return v;
}

// spline over value slots (df_new.c:705-728), recording the coordinates it reads
template <int D>
__device__ float tspline(const TProg &P, int spi, const float *vals, const int32_t *slots, uint32_t &used) {
    if constexpr (D == 0) {
        return __builtin_nanf("");                      // unreachable: the host bounds the depth
    } else {
        const McgSpline s = P.sp[spi];
        if (s.is_const) return s.value;
        const float in = vals[slots[s.coord] * kTWG];
        used |= 1u << s.coord;
        const float *loc = P.splf + s.loc0, *der = P.splf + s.loc0 + s.n;
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
            const float v = tspline<D - 1>(P, P.spc[s.val0 + base + q], vals, slots, used);
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

__device__ __forceinline__ float tgrad(const TIns &s, int coord) {             // df_new.c:674-694
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

// OR of the spline's used-coordinate mask into flags[chunk]: one atomic per wave when the wave's lanes
// share a chunk, else one per lane that read a coordinate
__device__ __forceinline__ void or_mask(uint32_t *flags, int c, bool active, uint32_t used) {
    const int c0 = __shfl(c, 0), c63 = __shfl(c, 63);
    if (c0 == c63) {
        uint32_t m = active ? used : 0u;
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) m |= __shfl_xor(m, d);
        if ((threadIdx.x & 63) == 0 && m) atomicOr(flags + c0, m);
    } else if (active && used) {
        atomicOr(flags + c, used);
    }
}

// FEAT (TF_* of mcg_terrain.h) selects the node kinds a pass holds: a pass without noises or splines (the
// block-level interpolation and arithmetic) compiles to a small kernel with few registers; SPD bounds the
// spline nesting.
template <int FEAT, int SPD>
__global__ __launch_bounds__(kTWG) void k_tpass(TProg P, TPass A, TBatch B) {
    extern __shared__ __align__(16) uint8_t lds[];
    GradEnt *G = (GradEnt *)lds;
    dev::init_grad(G);
    __syncthreads();
    float *vals = (float *)(lds + 256 * sizeof(GradEnt)) + threadIdx.x;
    // the host keeps n * nchunks below 2^31: 32-bit index arithmetic
    const uint32_t gid = blockIdx.x * kTWG + threadIdx.x;
    const int n = A.n;
    const bool active = gid < (uint32_t)n * (uint32_t)B.nchunks;
    const int c = active ? (int)(gid / (uint32_t)n) : 0;
    const int e = active ? (int)(gid - (uint32_t)c * (uint32_t)n) : 0;
    const TVol v = A.v;
    const int t = (int)((uint32_t)e / (uint32_t)v.sy), iy = e - t * v.sy;
    const int iz = (int)((uint32_t)t / (uint32_t)v.sx), ix = t - iz * v.sx;
    const int cx0 = B.cx0[c], cz0 = B.cz0[c];
    const int ox0 = tvx0(v, cx0), oz0 = tvz0(v, cz0);
    const int bx = ox0 + ix * v.dx, by = v.y0 + iy * v.dy, bz = oz0 + iz * v.dz;
#define V(x) vals[(x) * kTWG]
#pragma unroll 1
    for (int i = 0; i < A.nins; i++) {
        const TIns s = cload(A.ins + i);
        float r;
        switch (s.op) {
        case MCGK_CONST: case MCGK_CTX_ALPHA: case MCGK_CTX_OFFSET: r = s.f0; break;
        case MCGK_NOISE:
            if constexpr ((FEAT & TF_NOISE) != 0) r = vol_stack(P, G, s.ns, bx, by, bz, s.d0, s.d1, 1.0f);
            else r = __builtin_nanf("");
            break;
        case MCGK_SHIFT_B:                                                             // transposed volume
            if constexpr ((FEAT & TF_NOISE) != 0) r = vol_stack(P, G, s.ns, bz, bx, 0, s.d0, s.d0, 4.0f);
            else r = __builtin_nanf("");
            break;
        case MCGK_NOISE_XZ: case MCGK_NOISE_XYZ: {
            if constexpr ((FEAT & TF_NOISE) != 0) {
                const double nx = bx * s.d0 + (double)V(s.a);
                double ny = by * s.d1;
                if (s.op == MCGK_NOISE_XYZ) ny += (double)V(s.b);
                const double nz = bz * s.d0 + (double)V(s.c);
                r = pt_stack(P, G, s.ns, nx, ny, nz);
            } else {
                r = __builtin_nanf("");
            }
            break;
        }
        case MCGK_END:
            if constexpr ((FEAT & TF_NOISE) != 0) r = (end_height_j<false>(P.octs[s.oct].p, bx / 8, bz / 8) - 8.0f) / 128.0f;
            else r = __builtin_nanf("");
            break;
        case MCGK_DIST: r = dist_value(s.i3, (float)(s.i0 - bx), (float)(s.i1 - by), (float)(s.i2 - bz)); break;
        case MCGK_GRAD_CLAMP: case MCGK_GRAD_REPEAT: case MCGK_GRAD_MIRROR:
            r = tgrad(s, s.i0 == 0 ? bx : (s.i0 == 1 ? by : bz));
            break;
        case MCGK_ABS: r = fabsf(V(s.a)); break;
        case MCGK_SQUARE: { const float q = V(s.a); r = q * q; break; }
        case MCGK_CUBE: { const float q = V(s.a); r = q * q * q; break; }
        case MCGK_SQRT: r = (float)sqrt((double)V(s.a)); break;
        case MCGK_LEAKY: { const float q = V(s.a); r = q > 0.0f ? q : q * s.f0; break; }
        case MCGK_RECIP: r = 1.0f / V(s.a); break;
        case MCGK_NEG: r = -V(s.a); break;
        case MCGK_SQUEEZE: r = squeeze_f(V(s.a)); break;
        case MCGK_SIGN: r = jsignf(V(s.a)); break;
        case MCGK_CADD: r = V(s.a) + s.f0; break;
        case MCGK_CSUB: r = s.f0 - V(s.a); break;
        case MCGK_CMUL: r = V(s.a) * s.f0; break;
        case MCGK_CDIV: r = s.f0 / V(s.a); break;
        case MCGK_CMIN: { r = V(s.a); if (s.f0 < r) r = s.f0; break; }
        case MCGK_CMAX: { r = V(s.a); if (s.f0 > r) r = s.f0; break; }
        case MCGK_CLAMP: r = jclampf(V(s.a), s.f0, s.f1); break;
        case MCGK_RANGE_C: { const float q = V(s.a); r = (q >= s.f0 && q < s.f1) ? s.f2 : s.f3; break; }
        case MCGK_ADD: r = V(s.a) + V(s.b); break;
        case MCGK_SUB: r = V(s.a) + -V(s.b); break;
        case MCGK_MUL: r = V(s.a) * V(s.b); break;
        case MCGK_DIV: r = V(s.a) / V(s.b); break;
        case MCGK_MIN: { r = V(s.a); const float q = V(s.b); if (q < r) r = q; break; }
        case MCGK_MAX: { r = V(s.a); const float q = V(s.b); if (q > r) r = q; break; }
        case MCGK_LERP: case MCGK_LERP_CF: case MCGK_LERP_CS: {
            const float a = V(s.a);
            const float fv = s.op == MCGK_LERP_CF ? s.f0 : V(s.b), sv = s.op == MCGK_LERP_CS ? s.f0 : V(s.c);
            r = a == 0.0f ? fv : (a == 1.0f ? sv : jlerpf(a, fv, sv));
            break;
        }
        case MCGK_RANGE: { const float q = V(s.a); r = (!(q >= s.f0) || !(q < s.f1)) ? V(s.c) : V(s.b); break; }
        case MCGK_ISEL1: r = V(s.a) < s.f0 ? V(s.b) : V(s.c); break;
        case MCGK_ISEL: {
            const float q = V(s.a);
            int k = 0;
            for (; k < s.nthr; k++)
                if (q < (float)P.thr[s.thr0 + k]) break;
            if (k == s.nthr) k = s.narr - 1;
            r = V(A.arr[s.arr0 + k]);
            break;
        }
        case MCGK_SPLINE: {
            if constexpr ((FEAT & TF_SPLINE) != 0) {
                uint32_t used = 0;
                r = tspline<SPD>(P, s.sp, vals, A.arr + s.arr0, used);
                if (B.flags && s.flag >= 0) or_mask(B.flags + (int64_t)s.flag * B.flag_stride, c, active, used);
            } else {
                r = __builtin_nanf("");
            }
            break;
        }
        case TOP_LOAD: r = active ? B.arena[s.src + (int64_t)c * s.g0 + e] : 0.0f; break;
        case TOP_INTERP: {
            // cell-step interpolation (df_new.c:867-915) at one element of a unit-step volume; the corner
            // volume (g1 x g2 x g3, step cxz = i0, cy = i1) sits at src, g0 floats per chunk
            const int cxz = s.i0, cy = s.i1;
            int mnx, mny, mnz, cellx, celly, cellz;
            if (s.g4 >= 0) {                            // power-of-two cells: floor division by shifts
                const int lx = s.g4 & 255, ly = s.g4 >> 8;
                mnx = ox0 >> lx; mny = v.y0 >> ly; mnz = oz0 >> lx;
                cellx = (bx >> lx) - mnx; celly = (by >> ly) - mny; cellz = (bz >> lx) - mnz;
            } else {
                mnx = jfloordiv(ox0, cxz); mny = jfloordiv(v.y0, cy); mnz = jfloordiv(oz0, cxz);
                cellx = jfloordiv(bx, cxz) - mnx; celly = jfloordiv(by, cy) - mny; cellz = jfloordiv(bz, cxz) - mnz;
            }
            const int cox = (mnx + cellx) * cxz - ox0, coy = (mny + celly) * cy - v.y0, coz = (mnz + cellz) * cxz - oz0;
            const int xx = ix - cox, yy = iy - coy, zz = iz - coz;
            const int ystart = coy < 0 ? -coy : 0;
            const int cvx = s.g1, cvy = s.g2, cvz = s.g3;
            const int nx2 = cellx + 1 < cvx - 1 ? cellx + 1 : cvx - 1;
            const int ny2 = celly + 1 < cvy - 1 ? celly + 1 : cvy - 1;
            const int nz2 = cellz + 1 < cvz - 1 ? cellz + 1 : cvz - 1;
            const float *cb = B.arena + s.src + (int64_t)c * s.g0;
            auto at = [&](int x, int y, int z) { return active ? cb[y + (x + z * cvx) * cvy] : 0.0f; };
            const float v000 = at(cellx, celly, cellz), v100 = at(nx2, celly, cellz);
            const float v001 = at(cellx, celly, nz2), v101 = at(nx2, celly, nz2);
            const float v010 = at(cellx, ny2, cellz), v110 = at(nx2, ny2, cellz);
            const float v011 = at(cellx, ny2, nz2), v111 = at(nx2, ny2, nz2);
            const float az = (float)zz * s.f0;
            const float v00 = jlerpf(az, v000, v001), v01 = jlerpf(az, v010, v011);
            const float v10 = jlerpf(az, v100, v101), v11 = jlerpf(az, v110, v111);
            const float ax = (float)xx * s.f0;
            const float a0 = jlerpf(ax, v00, v10), a1 = jlerpf(ax, v01, v11);
            const float step = (a1 - a0) * s.f1;
            float value = a0 + step * (float)ystart;
#pragma unroll 1
            for (int k = ystart; k < yy; k++) value += step;
            r = value;
            break;
        }
        case TOP_SLICE: {
            // broadcast from the volume with axis g1 fixed (0 x, 1 y, 2 z, 3 x and z), g0 floats per chunk
            int j;
            switch (s.g1) {
            case 0: j = iy + iz * v.sy; break;
            case 1: j = ix + iz * v.sx; break;
            case 2: j = iy + ix * v.sy; break;
            default: j = iy; break;
            }
            r = active ? B.arena[s.src + (int64_t)c * s.g0 + j] : 0.0f;
            break;
        }
        default: r = __builtin_nanf(""); break;          // the host refuses other kinds
        }
        V(s.dst) = r;
        if (s.store >= 0 && active) B.arena[s.store + (int64_t)c * n + e] = r;
    }
#undef V
}

hipError_t launch_tpass(const TProg &P, const TPass &A, const TBatch &B, hipStream_t st) {
    const int64_t items = (int64_t)A.n * B.nchunks;
    if (items <= 0) return hipSuccess;
    const int lds = 256 * (int)sizeof(GradEnt) + A.nslot * kTWG * 4;
    const dim3 grid((unsigned)((items + kTWG - 1) / kTWG));
    if (!(A.feat & TF_SPLINE)) {
        if (A.feat & TF_NOISE) hipLaunchKernelGGL((k_tpass<TF_NOISE, 1>), grid, dim3(kTWG), lds, st, P, A, B);
        else hipLaunchKernelGGL((k_tpass<0, 1>), grid, dim3(kTWG), lds, st, P, A, B);
    } else if (A.spd <= kShallowSpline) {
        hipLaunchKernelGGL((k_tpass<TF_NOISE | TF_SPLINE, kShallowSpline>), grid, dim3(kTWG), lds, st, P, A, B);
    } else {
        hipLaunchKernelGGL((k_tpass<TF_NOISE | TF_SPLINE, kSplineDepth>), grid, dim3(kTWG), lds, st, P, A, B);
    }
    return hipGetLastError();
}

// element-wise unary node of the volume path (df_new.c:979-1033)
__device__ __forceinline__ float tail_op(int op, float q, float f0, float f1, float f2, float f3) {
    switch (op) {
    case MCGK_ABS: return fabsf(q);
    case MCGK_SQUARE: return q * q;
    case MCGK_CUBE: return q * q * q;
    case MCGK_SQRT: return (float)sqrt((double)q);
    case MCGK_LEAKY: return q > 0.0f ? q : q * f0;
    case MCGK_RECIP: return 1.0f / q;
    case MCGK_NEG: return -q;
    case MCGK_SQUEEZE: return squeeze_f(q);
    case MCGK_SIGN: return jsignf(q);
    case MCGK_CADD: return q + f0;
    case MCGK_CSUB: return f0 - q;
    case MCGK_CMUL: return q * f0;
    case MCGK_CDIV: return f0 / q;
    case MCGK_CMIN: return f0 < q ? f0 : q;
    case MCGK_CMAX: return f0 > q ? f0 : q;
    case MCGK_CLAMP: return jclampf(q, f0, f1);
    default: return (q >= f0 && q < f1) ? f2 : f3;     // MCGK_RANGE_C
    }
}

__global__ __launch_bounds__(kTWG) void k_tinterp(TInterp I, TVol v, TBatch B) {
    const int cy = 1 << I.ly, ny = v.sy >> I.ly;
    const uint32_t per = (uint32_t)(v.sx * v.sz * ny);
    const uint32_t gid = blockIdx.x * kTWG + threadIdx.x;
    if (gid >= per * (uint32_t)B.nchunks) return;      // no barrier in this kernel
    const int c = (int)(gid / per), r = (int)(gid - (uint32_t)c * per);
    const int run = r % ny, col = r / ny, ix = col % v.sx, iz = col / v.sx;
    const int cellx = ix >> I.lx, cellz = iz >> I.lx, celly = run;
    const int xx = ix - (cellx << I.lx), zz = iz - (cellz << I.lx);
    const int nx2 = cellx + 1 < I.cvx - 1 ? cellx + 1 : I.cvx - 1;
    const int ny2 = celly + 1 < I.cvy - 1 ? celly + 1 : I.cvy - 1;
    const int nz2 = cellz + 1 < I.cvz - 1 ? cellz + 1 : I.cvz - 1;
    const float *cb = B.arena + I.src + (int64_t)c * I.ncv;
    auto at = [&](int x, int y, int z) { return cb[y + (x + z * I.cvx) * I.cvy]; };
    const float v000 = at(cellx, celly, cellz), v100 = at(nx2, celly, cellz);
    const float v001 = at(cellx, celly, nz2), v101 = at(nx2, celly, nz2);
    const float v010 = at(cellx, ny2, cellz), v110 = at(nx2, ny2, cellz);
    const float v011 = at(cellx, ny2, nz2), v111 = at(nx2, ny2, nz2);
    const float az = (float)zz * I.f0;
    const float v00 = jlerpf(az, v000, v001), v01 = jlerpf(az, v010, v011);
    const float v10 = jlerpf(az, v100, v101), v11 = jlerpf(az, v110, v111);
    const float ax = (float)xx * I.f0;
    const float a0 = jlerpf(ax, v00, v10), a1 = jlerpf(ax, v01, v11);
    const float step = (a1 - a0) * I.f1;
    float value = a0 + step * 0.0f;
    float *out = B.arena + I.store + (int64_t)c * v.size() + (int64_t)(ix + iz * v.sx) * v.sy + (int64_t)run * cy;
    // runs of four blocks leave as one 16-byte store (the run starts 16-byte aligned: arena offsets are
    // multiples of 64 floats and cy, sy of 4)
    for (int k = 0; k < cy; k += 4) {
        float q[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            float w = value;
            for (int t = 0; t < I.ntail; t++) w = tail_op(I.op[t], w, I.k0[t], I.k1[t], I.k2[t], I.k3[t]);
            q[u] = w;
            value += step;
        }
        *reinterpret_cast<float4 *>(out + k) = make_float4(q[0], q[1], q[2], q[3]);
    }
}

hipError_t launch_tinterp(const TInterp &I, const TVol &v, const TBatch &B, hipStream_t st) {
    const int64_t items = (int64_t)v.sx * v.sz * (v.sy >> I.ly) * B.nchunks;
    if (items <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tinterp, dim3((unsigned)((items + kTWG - 1) / kTWG)), dim3(kTWG), 0, st, I, v, B);
    return hipGetLastError();
}

// Mth.getSeed(x, y, z) (engine/mc_rng.h:124-128) and the positional Xoroshiro128++ floats
__device__ __forceinline__ int64_t get_seed_xyz(int32_t x, int32_t y, int32_t z) {
    int64_t l = (int64_t)(int32_t)((uint32_t)x * 3129871u) ^ ((int64_t)z * 116129781LL) ^ (int64_t)y;
    l = (int64_t)((uint64_t)l * (uint64_t)l * 42317861ULL + (uint64_t)l * 11ULL);
    return l >> 16;
}
struct Xo {
    uint64_t lo, hi;
    __device__ float next_float() {
        const uint64_t s0 = lo;
        uint64_t s1 = hi;
        const uint64_t res = ((s0 + s1) << 17 | (s0 + s1) >> 47) + s0;
        s1 ^= s0;
        lo = (s0 << 49 | s0 >> 15) ^ s1 ^ (s1 << 21);
        hi = s1 << 28 | s1 >> 36;
        return (float)(res >> 40) * 5.9604645E-8f;
    }
};

__global__ __launch_bounds__(kTWG) void k_tveins(PWorld G, TVeins V, TBatch B, uint16_t *out) {
    extern __shared__ __align__(16) uint8_t lds[];
    dev::init_grad((GradEnt *)lds);
    __syncthreads();
    float *vals = (float *)(lds + 256 * sizeof(GradEnt)) + threadIdx.x;
    const int nd = 256 * V.nh;
    const int64_t gid = (int64_t)blockIdx.x * kTWG + threadIdx.x;
    if (gid >= (int64_t)nd * B.nchunks) return;        // no barrier below
    const int c = (int)(gid / nd), e = (int)(gid - (int64_t)c * nd);
    const int iy = e % V.nh, t = e / V.nh, ix = t & 15, iz = t >> 4;
    const int bx = B.cx0[c] + ix, by = V.nmin + iy, bz = B.cz0[c] + iz;
    uint16_t result = 0xFFFF;
    for (int r = 0; r < V.n; r++) {
        const float d = B.arena[V.dens[r] + (int64_t)c * nd + e];
        if (d <= 0.0f) continue;
        Xo x{(uint64_t)get_seed_xyz(bx, by, bz) ^ V.lo, V.hi};
        if ((x.lo | x.hi) == 0) { x.lo = 0x9E3779B97F4A7C15ULL; x.hi = 0x6A09E667F3BCC909ULL; }
        if (x.next_float() > d) continue;
        bool pass = x.next_float() < B.arena[V.rich[r] + (int64_t)c * nd + e];
        if (pass) {
            eval_new<false, kSplineDepth>(G, lds, vals, bx, by, bz, false);
            pass = vals[V.gap_slot[r] * kWG] < 0.0f;
        }
        int outst;
        if (pass) outst = x.next_float() < V.raw_chance[r] ? V.raw[r] : V.ore[r];
        else outst = V.filler[r];
        result = (uint16_t)outst;
        break;
    }
    out[(int64_t)c * nd + (iy * 16 + iz) * 16 + ix] = result;
}

hipError_t launch_tveins(const PWorld &G, int lds, const TVeins &V, const TBatch &B, uint16_t *out, hipStream_t st) {
    const int64_t items = (int64_t)256 * V.nh * B.nchunks;
    if (items <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tveins, dim3((unsigned)((items + kTWG - 1) / kTWG)), dim3(kTWG), lds, st, G, V, B, out);
    return hipGetLastError();
}

__global__ __launch_bounds__(kTWG) void k_tgather(const TGather *items, const float *arena, float *dst) {
    const TGather g = items[blockIdx.y];
    for (int i = blockIdx.x * kTWG + threadIdx.x; i < g.n; i += gridDim.x * kTWG) dst[g.dst + i] = arena[g.src + i];
}

hipError_t launch_tgather(const TGather *items, int nitems, const float *arena, float *dst, hipStream_t st) {
    if (nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tgather, dim3(4, (unsigned)nitems), dim3(kTWG), 0, st, items, arena, dst);
    return hipGetLastError();
}

}  // namespace plug
}  // namespace mcb
