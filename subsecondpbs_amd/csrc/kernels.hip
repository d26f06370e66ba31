// gfx950 kernels of the pencil-beam final-dose path.
//
//   k_trace ............ BEV fan-ray trace: explicit trilinear CT gather, density and
//                        compensated WEPL prefix per ray (bev_ray_tracing.cu:65-159)
//   k_trace_reduce ..... per-step min WEPL and beam entry/exit steps
//                        (complete_superposition.cu:109-177, wrapper:1990-2023)
//   k_beam_setup ....... distal cut-offs and entry sigmas on the device (wrapper:3651-3736,
//                        4308-4329), so a beam needs no host round trip
//   k_conv_x / k_conv_y  separable erf gather-resample spot lattice -> rays
//                        (convolution.cu:198-262, 477-513)
//   k_subspot_cpb ...... subspot -> CPB 2D erf cell integrals as a gather, no atomics
//                        (convolution.cu:95-152)
//   k_cpb_to_rays ...... area-overlap resample (convolution.cu:269-339)
//   k_fill ............. IDD and sigma per ray and layer (idd_sigma.cu:46-259)
//   k_layer_finalize ... layer passive steps, layer groups (layers sharing a passive step
//                        share a transfer box), group boxes and tail planes
//                        (wrapper:4950-4996, 5343-5395)
//   k_tiles / k_bins ... 32x8 tile radii, radius histogram, batched launch radii
//                        (complete_superposition.cu:197-255, 400-490)
//   k_items ............ per (step, tile) work item: max launch radius over live layers
//   k_superpose<RM> .... register-blocked Gaussian superposition of one (step, tile) over
//                        all layers into private output blocks (main + group tails; no
//                        atomics)
//   k_merge ............ deterministic sum of overlapping blocks into the main BEV volume
//                        and the group tail planes
//   k_transfer ......... divergent BEV -> dose-grid trilinear gather, one owner per dose
//                        column (prim_transf_kernel.cu:11-44); voxels inside a group's
//                        transfer box also sample that group's tail planes
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>
#include <functional>
#include <type_traits>

#include "device.h"
#include "kernels.h"

namespace pbs {

#define PBS_INF __builtin_inff()

__device__ __forceinline__ float lut1d_clamp(const float *t, int n, float xb) {
    const float f = floorf(xb);
    const float a = xb - f;
    int i0 = int(f), i1 = i0 + 1;
    i0 = min(max(i0, 0), n - 1);
    i1 = min(max(i1, 0), n - 1);
    return (1.0f - a) * t[i0] + a * t[i1];
}

__device__ __forceinline__ float lut2d_clamp(const float *__restrict__ t, int nx, int ny, float xb, float yb) {
    const float fx = floorf(xb), fy = floorf(yb);
    const float ax = xb - fx, ay = yb - fy;
    int x0 = int(fx), y0 = int(fy);
    int x1 = min(max(x0 + 1, 0), nx - 1), y1 = min(max(y0 + 1, 0), ny - 1);
    x0 = min(max(x0, 0), nx - 1);
    y0 = min(max(y0, 0), ny - 1);
    const float v0 = (1.0f - ax) * t[y0 * nx + x0] + ax * t[y0 * nx + x1];
    const float v1 = (1.0f - ax) * t[y1 * nx + x0] + ax * t[y1 * nx + x1];
    return (1.0f - ay) * v0 + ay * v1;
}

template <typename T>
__device__ __forceinline__ float trilinear_border(const T *__restrict__ v, int nx, int ny, int nz, float x, float y,
                                                  float z) {
    const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    const int x0 = int(fx), y0 = int(fy), z0 = int(fz);
    const float tx = x - fx, ty = y - fy, tz = z - fz;
    float c[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int xi = x0 + (q & 1), yi = y0 + ((q >> 1) & 1), zi = z0 + (q >> 2);
        const bool in = unsigned(xi) < unsigned(nx) && unsigned(yi) < unsigned(ny) && unsigned(zi) < unsigned(nz);
        c[q] = in ? float(v[(size_t(zi) * ny + yi) * nx + xi]) : 0.0f;
    }
    const float c00 = c[0] + (c[1] - c[0]) * tx;
    const float c10 = c[2] + (c[3] - c[2]) * tx;
    const float c01 = c[4] + (c[5] - c[4]) * tx;
    const float c11 = c[6] + (c[7] - c[6]) * tx;
    const float c0 = c00 + (c10 - c00) * ty;
    const float c1 = c01 + (c11 - c01) * ty;
    return c0 + (c1 - c0) * tz;
}

// Trilinear sample of a stack of planes with border 0, where only planes [zlo, zhi) hold data
// (the others read as 0 and are never written).
__device__ __forceinline__ float trilinear_planes(const float *__restrict__ v, int nx, int ny, int zlo, int zhi,
                                                  float x, float y, float z) {
    const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    const int x0 = int(fx), y0 = int(fy), z0 = int(fz);
    const float tx = x - fx, ty = y - fy, tz = z - fz;
    float c[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int xi = x0 + (q & 1), yi = y0 + ((q >> 1) & 1), zi = z0 + (q >> 2);
        const bool in = unsigned(xi) < unsigned(nx) && unsigned(yi) < unsigned(ny) && zi >= zlo && zi < zhi;
        c[q] = in ? v[(size_t(zi) * ny + yi) * nx + xi] : 0.0f;
    }
    const float c00 = c[0] + (c[1] - c[0]) * tx;
    const float c10 = c[2] + (c[3] - c[2]) * tx;
    const float c01 = c[4] + (c[5] - c[4]) * tx;
    const float c11 = c[6] + (c[7] - c[6]) * tx;
    const float c0 = c00 + (c10 - c00) * ty;
    const float c1 = c01 + (c11 - c01) * ty;
    return c0 + (c1 - c0) * tz;
}

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
// The same minimum as a uniform value, through DPP lane moves within each row of 16 lanes (the
// quad swaps, then the half-row and row mirrors) and four lane reads across the rows: no LDS
// round trips. All 64 lanes must be active. fminf ignores NaN, and a minimum does not depend on
// the order it is taken in.
// The minimum over each row of 16 lanes, in every lane of the row (rows are whole: their lanes
// are all active or all inactive).
__device__ __forceinline__ float row_min(float v) {
    auto dmin = [](float x, auto ctrl) {
        const int y = __builtin_amdgcn_update_dpp(__float_as_int(x), __float_as_int(x), decltype(ctrl)::value,
                                                  0xf, 0xf, false);
        return fminf(x, __int_as_float(y));
    };
    v = dmin(v, std::integral_constant<int, 0xb1>{});   // quad_perm [1, 0, 3, 2]
    v = dmin(v, std::integral_constant<int, 0x4e>{});   // quad_perm [2, 3, 0, 1]
    v = dmin(v, std::integral_constant<int, 0x141>{});  // row_half_mirror
    return dmin(v, std::integral_constant<int, 0x140>{});  // row_mirror
}
__device__ __forceinline__ float wave_min_uniform(float v) {
    v = row_min(v);
    const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
    const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
    const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
    const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
    return fminf(fminf(r0, r1), fminf(r2, r3));
}
// Integer maximum / minimum over the wave as uniform values, the same way (all 64 lanes active).
template <bool MAX>
__device__ __forceinline__ int wave_ext_uniform_i(int v) {
    auto dext = [](int x, auto ctrl) {
        const int y = __builtin_amdgcn_update_dpp(x, x, decltype(ctrl)::value, 0xf, 0xf, false);
        return MAX ? max(x, y) : min(x, y);
    };
    v = dext(v, std::integral_constant<int, 0xb1>{});
    v = dext(v, std::integral_constant<int, 0x4e>{});
    v = dext(v, std::integral_constant<int, 0x141>{});
    v = dext(v, std::integral_constant<int, 0x140>{});
    const int r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
    const int r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
    return MAX ? max(max(r0, r1), max(r2, r3)) : min(min(r0, r1), min(r2, r3));
}
__device__ __forceinline__ int wave_max_uniform_i(int v) { return wave_ext_uniform_i<true>(v); }
__device__ __forceinline__ int wave_min_uniform_i(int v) { return wave_ext_uniform_i<false>(v); }
__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// findFirstLargerOrdered (include/utils/vector_find.h:62-81) on a device array.
__device__ int first_larger_ordered(const float *v, int n, float value) {
    int upper = n - 1, lower = 0;
    if (v[n - 1] <= value) return upper;
    if (v[0] > value) return 0;
    while (upper - lower > 1) {
        const int pivot = (upper + lower) / 2;
        if (v[pivot] <= value) lower = pivot; else upper = pivot;
    }
    return lower + 1;
}

// ---------------------------------------------------------------------------------------
// trace
// ---------------------------------------------------------------------------------------
// The trace runs as (ray, 8-step segment) threads so that every ray's K steps are spread
// over many waves: pass 1 samples the CT and keeps segment-local prefixes, pass 2 turns the
// segment totals into carries and the entry / exit steps, pass 3 adds the carries.
constexpr int kTraceSeg = 8;

// A ray's start and per-step increment in CT index space, and its step length. The increment
// is the exact derivative of the divergent fan point (not a difference of two rounded points,
// whose error would grow with the step index).
struct RayGeom {
    float p0[3], inc[3], stepLen;
};
__device__ __forceinline__ RayGeom ray_geom(const FanGeom &g, int ray) {
    RayGeom rg;
    const int i = ray % g.rx, j = ray / g.rx;
    const float fx = g.fc[0] + float(i) * g.fd[0];
    const float fy = g.fc[1] + float(j) * g.fd[1];
    const float fz = g.fc[2];
    const float gx = fx * (1.0f - fz / g.dist[0]);
    const float gy = fy * (1.0f - fz / g.dist[1]);
    const float dgx = -fx * (g.fd[2] / g.dist[0]);
    const float dgy = -fy * (g.fd[2] / g.dist[1]);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        rg.p0[r] = g.g2im[3 * r] * gx + g.g2im[3 * r + 1] * gy + g.g2im[3 * r + 2] * fz + g.g2imOff[r];
        rg.inc[r] = g.g2im[3 * r] * dgx + g.g2im[3 * r + 1] * dgy + g.g2im[3 * r + 2] * g.fd[2];
    }
    const float dx = fx / g.dist[0], dy = fy / g.dist[1];
    rg.stepLen = fabsf(g.fd[2]) * sqrtf(1.0f + dx * dx + dy * dy);
    return rg;
}

// Running sums of a segment: the Kahan-compensated WEPL prefix (cum, comp), the HU prefix, and
// the last step with HU+1000 > 150.
struct TraceAcc {
    float cum, comp, cumHu;
    int lastIn = -1;
};

// Step k of a ray (bev_ray_tracing.cu:65-159): explicit trilinear CT sample (texel centre -1/2,
// border 0), the running sums advanced by it; returns the step's density. Both trace forms
// take their steps through this one function.
__device__ __forceinline__ float trace_step(const FanGeom &g, const float *__restrict__ ct, const LutDev &lut,
                                            const float *s_dens, const float *s_sp, const RayGeom &rg, int k,
                                            TraceAcc &a) {
    const float fk = float(k);
    const float hu = trilinear_border(ct, g.ctDims[0], g.ctDims[1], g.ctDims[2], rg.p0[0] + fk * rg.inc[0],
                                      rg.p0[1] + fk * rg.inc[1], rg.p0[2] + fk * rg.inc[2]);
    a.cumHu += hu;
    const float dens = lut1d_clamp(s_dens, lut.nDensity, hu * lut.densityScale);
    const float y = rg.stepLen * lut1d_clamp(s_sp, lut.nSp, hu * lut.spScale) - a.comp;
    const float t = a.cum + y;
    a.comp = (t - a.cum) - y;
    a.cum = t;
    if (hu > 150.0f) a.lastIn = k;
    return dens;
}

// Pass 1 (bev_ray_tracing.cu:65-159 per segment): explicit trilinear CT sample at fan step k
// (texel centre -1/2, border 0), density and SP from the LUTs in LDS, the segment's Kahan
// prefix of stepLen * SP (written to wepl), the segment's HU prefix, its totals and its last
// step with HU+1000 > 150.
__device__ __forceinline__ void trace_seg_body(const uint3 blockIdx, const uint3 gridDim, FanGeom g, const float *__restrict__ ct, LutDev lut,
                                                   float *__restrict__ density, float *__restrict__ wepl,
                                                   float *__restrict__ huPrefix, float2 *__restrict__ segSp,
                                                   float *__restrict__ segHu, int *__restrict__ segLast,
                                                   BeamScalars *__restrict__ bs) {
    extern __shared__ float s_lut[];  // density LUT, then stopping-power LUT
    for (int i = threadIdx.x; i < lut.nDensity; i += blockDim.x) s_lut[i] = lut.density[i];
    for (int i = threadIdx.x; i < lut.nSp; i += blockDim.x) s_lut[lut.nDensity + i] = lut.sp[i];
    __syncthreads();
    const int nray = g.rx * g.ry;
    const int ray = blockIdx.x * blockDim.x + threadIdx.x;
    const int seg = blockIdx.y;
    if (ray >= nray) return;
    if (ray == 0 && seg == 0) {
        // the carry pass reduces the beam's entry / exit steps into these
        bs->beamFirstInside = INT_MAX;
        bs->beamFirstOutside = INT_MIN;
    }
    const float *s_dens = s_lut, *s_sp = s_lut + lut.nDensity;
    const RayGeom rg = ray_geom(g, ray);
    const int k0 = seg * kTraceSeg, k1 = min(g.K, k0 + kTraceSeg);
    TraceAcc a{};
    size_t idx = size_t(k0) * nray + ray;
#pragma unroll 4
    for (int k = k0; k < k1; ++k) {
        density[idx] = trace_step(g, ct, lut, s_dens, s_sp, rg, k, a);
        wepl[idx] = a.cum;
        huPrefix[idx] = a.cumHu;
        idx += nray;
    }
    const float cum = a.cum, comp = a.comp, cumHu = a.cumHu;
    const int lastIn = a.lastIn;
    const size_t o = size_t(seg) * nray + ray;
    segSp[o] = make_float2(cum, comp);
    segHu[o] = cumHu;
    segLast[o] = lastIn;
}
struct TraceSegK {
    static constexpr int NT = 256;
    struct Args {
        FanGeom g;
        const float *ct;
        LutDev lut;
        float *density;
        float *wepl;
        float *huPrefix;
        float2 *segSp;
        float *segHu;
        int *segLast;
        BeamScalars *bs;
    };
    static __device__ __forceinline__ void run(const Args &a, const uint3 b, const uint3 d) { trace_seg_body(b, d, a.g, a.ct, a.lut, a.density, a.wepl, a.huPrefix, a.segSp, a.segHu, a.segLast, a.bs); }
};


// Pass 2, thread per ray: segment carries (double), entry and exit steps. The cumulative HU
// never decreases (HU+1000 >= 0), so the reference's flag loop (bev_ray_tracing.cu:141-158)
// reduces to: firstInside = max(#steps with cumulative HU < 150, 1); firstOutside =
// last step with HU+1000 > 150 (0 if none) + 1.
__device__ __forceinline__ void trace_carry_body(const uint3 blockIdx, const uint3 gridDim, FanGeom g, const float *__restrict__ huPrefix,
                                                     const float2 *__restrict__ segSp,
                                                     const float *__restrict__ segHu, const int *__restrict__ segLast,
                                                     double *__restrict__ segCarry, int *__restrict__ firstInside,
                                                     int *__restrict__ firstOutside, unsigned *__restrict__ weplMinBits,
                                                     BeamScalars *__restrict__ bs) {
    const int nray = g.rx * g.ry;
    const int ray = blockIdx.x * blockDim.x + threadIdx.x;
    // the per-step minimum that k_trace_fix reduces into starts at 1e20 (Macro.cuh INF)
    if (ray < g.K) weplMinBits[ray] = __float_as_uint(1.0e20f);
    // whole waves past the last ray leave; lanes past it take neutral values in the reductions
    if (ray - int(threadIdx.x & 63) >= nray) return;
    int fiv = INT_MAX, fov = INT_MIN;
    if (ray < nray) {
        const int S = (g.K + kTraceSeg - 1) / kTraceSeg;
        double carrySp = 0.0, carryHu = 0.0;
        int count = -1, last = -1;
        for (int sg = 0; sg < S; ++sg) {
            const size_t o = size_t(sg) * nray + ray;
            segCarry[o] = carrySp;
            const float2 sp = segSp[o];
            const double hu = double(segHu[o]);
            if (count < 0 && carryHu + hu >= 150.0) {
                // the threshold is crossed in this segment: find the step
                const int k0 = sg * kTraceSeg, k1 = min(g.K, k0 + kTraceSeg);
                count = k1;
                for (int k = k0; k < k1; ++k)
                    if (carryHu + double(huPrefix[size_t(k) * nray + ray]) >= 150.0) { count = k; break; }
            }
            carrySp += double(sp.x) - double(sp.y);
            carryHu += hu;
            last = max(last, segLast[o]);
        }
        if (count < 0) count = g.K;
        fiv = max(count, 1);
        fov = max(last, 0) + 1;
        firstInside[ray] = fiv;
        firstOutside[ray] = fov;
    }
    // beamFirstInside / beamFirstOutside (k_trace_seg set them to INT_MAX / INT_MIN)
    const int mn = wave_min_i(fiv), mx = wave_max_i(fov);
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&bs->beamFirstInside, mn);
        atomicMax(&bs->beamFirstOutside, mx);
    }
}
struct TraceCarryK {
    static constexpr int NT = 256;
    struct Args {
        FanGeom g;
        const float *huPrefix;
        const float2 *segSp;
        const float *segHu;
        const int *segLast;
        double *segCarry;
        int *firstInside;
        int *firstOutside;
        unsigned *weplMinBits;
        BeamScalars *bs;
    };
    static __device__ __forceinline__ void run(const Args &a, const uint3 b, const uint3 d) { trace_carry_body(b, d, a.g, a.huPrefix, a.segSp, a.segHu, a.segLast, a.segCarry, a.firstInside, a.firstOutside, a.weplMinBits, a.bs); }
};


// Pass 3, thread per (step, ray): the segment's carry added to its local WEPL prefix (segment 0
// has none), and the step's minimum WEPL over the rays that leave material (fo > 0), the
// weplMin of the wrapper's per-beam reduction (wrapper:3651-3700): one atomic per block on the
// float's bits (WEPL >= 0, so the bit patterns order as the values).
__device__ __forceinline__ void trace_fix_body(const uint3 blockIdx, const uint3 gridDim, FanGeom g, const double *__restrict__ segCarry,
                                                   float *__restrict__ wepl, const int *__restrict__ fo,
                                                   unsigned *__restrict__ weplMinBits) {
    __shared__ float s_m[4];
    const int nray = g.rx * g.ry;
    const int ray = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    float m = 1.0e20f;
    if (ray < nray) {
        const size_t idx = size_t(k) * nray + ray;
        float v = wepl[idx];
        if (k >= kTraceSeg) {
            v = float(segCarry[size_t(k / kTraceSeg) * nray + ray] + double(v));
            wepl[idx] = v;
        }
        if (fo[ray] > 0) m = v + 0.0f;  // -0 -> +0
    }
    m = wave_min(m);
    if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float bm = fminf(fminf(s_m[0], s_m[1]), fminf(s_m[2], s_m[3]));
        if (bm < 1.0e20f) atomicMin(&weplMinBits[k], __float_as_uint(bm));
    }
}
struct TraceFixK {
    static constexpr int NT = 256;
    struct Args {
        FanGeom g;
        const double *segCarry;
        float *wepl;
        const int *fo;
        unsigned *weplMinBits;
    };
    static __device__ __forceinline__ void run(const Args &a, const uint3 b, const uint3 d) { trace_fix_body(b, d, a.g, a.segCarry, a.wepl, a.fo, a.weplMinBits); }
};


// The trace in one pass, for beams of up to 16 x CH segments (K <= 128 CH steps): a work-group
// of 256 threads holds 16 rays, 16 threads per ray, each thread a run of the ray's 8-step
// segments [s0, s1) (balanced: one or two per thread for K = 129). The segments are the ones of
// the three-pass form and each runs the same steps (trace_step); their totals meet in LDS, where
// each thread sums the carries of the segments before its own in segment order, in double, as
// the carry pass does (the two forms agree up to the compiler's choice of fused multiply-adds
// in their inlined copies). The beam's reductions leave as per-work-group partials (the
// per-step minimum over the group's rays that leave material, its least entry and greatest
// exit step), which the upload launch reduces: no atomics, and nothing to initialise first.
constexpr int kTraceRays = 16, kTraceThreadsPerRay = 16;

template <int CH>
__device__ __forceinline__ void trace_body(const uint3 blockIdx, FanGeom g, const float *__restrict__ ct, LutDev lut,
                                           float *__restrict__ density, float *__restrict__ wepl, int *__restrict__ fi,
                                           int *__restrict__ fo, float *__restrict__ partMin, int partPitch,
                                           int *__restrict__ partFi, int *__restrict__ partFo) {
    extern __shared__ float s_dyn[];
    const int nLut = lut.nDensity + lut.nSp;
    const int S = (g.K + kTraceSeg - 1) / kTraceSeg;
    float *s_dens = s_dyn, *s_sp = s_dyn + lut.nDensity;
    float2 *s_seg = reinterpret_cast<float2 *>(s_dyn + (nLut + 3) / 4 * 4);  // [S][rays] (cum, comp)
    float *s_hu = reinterpret_cast<float *>(s_seg + S * kTraceRays);       // [S][rays]
    int *s_last = reinterpret_cast<int *>(s_hu + S * kTraceRays);          // [S][rays]
    int *s_cnt = s_last + S * kTraceRays;                                   // [rays] entry step
    for (int i = threadIdx.x; i < lut.nDensity; i += blockDim.x) s_dens[i] = lut.density[i];
    for (int i = threadIdx.x; i < lut.nSp; i += blockDim.x) s_sp[i] = lut.sp[i];
    const int rl = threadIdx.x & (kTraceRays - 1), t = threadIdx.x / kTraceRays;
    const int nray = g.rx * g.ry;
    const int ray = blockIdx.x * kTraceRays + rl;
    const bool rayIn = ray < nray;
    const int s0 = t * S / kTraceThreadsPerRay, s1 = (t + 1) * S / kTraceThreadsPerRay;
    if (t == 0) s_cnt[rl] = g.K;  // no threshold crossing: K
    __syncthreads();
    float wv[CH * kTraceSeg], hv[CH * kTraceSeg];
    if (rayIn) {
        const RayGeom rg = ray_geom(g, ray);
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int sg = s0 + c;
            if (sg >= s1) break;
            TraceAcc a{};
#pragma unroll
            for (int j = 0; j < kTraceSeg; ++j) {
                const int k = sg * kTraceSeg + j;
                if (k >= g.K) break;
                density[size_t(k) * nray + ray] = trace_step(g, ct, lut, s_dens, s_sp, rg, k, a);
                wv[c * kTraceSeg + j] = a.cum;
                hv[c * kTraceSeg + j] = a.cumHu;
            }
            s_seg[sg * kTraceRays + rl] = make_float2(a.cum, a.comp);
            s_hu[sg * kTraceRays + rl] = a.cumHu;
            s_last[sg * kTraceRays + rl] = a.lastIn;
        }
    }
    __syncthreads();
    int fov = INT_MIN;
    if (rayIn) {
        // the carries of the segments before s0, summed in segment order
        double carrySp = 0.0, carryHu = 0.0;
        int last = -1;
        for (int sg = 0; sg < S; ++sg) {
            if (sg < s0) {
                const float2 sp = s_seg[sg * kTraceRays + rl];
                carrySp += double(sp.x) - double(sp.y);
                carryHu += double(s_hu[sg * kTraceRays + rl]);
            }
            last = max(last, s_last[sg * kTraceRays + rl]);
        }
        fov = max(last, 0) + 1;
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int sg = s0 + c;
            if (sg >= s1) break;
            const double hu = double(s_hu[sg * kTraceRays + rl]);
            const int k0 = sg * kTraceSeg, k1 = min(g.K, k0 + kTraceSeg);
            // the cumulative HU first reaches 150 in this segment: its step (the running HU sum
            // never decreases, so one segment at most sees the crossing)
            if (carryHu < 150.0 && carryHu + hu >= 150.0) {
                int count = k1;
#pragma unroll
                for (int j = kTraceSeg - 1; j >= 0; --j)
                    if (k0 + j < k1 && carryHu + double(hv[c * kTraceSeg + j]) >= 150.0) count = k0 + j;
                s_cnt[rl] = count;
            }
#pragma unroll
            for (int j = 0; j < kTraceSeg; ++j) {
                const int k = k0 + j;
                if (k >= k1) break;
                const float v = sg > 0 ? float(carrySp + double(wv[c * kTraceSeg + j])) : wv[c * kTraceSeg + j];
                wepl[size_t(k) * nray + ray] = v;
                wv[c * kTraceSeg + j] = v;
            }
            const float2 sp = s_seg[sg * kTraceRays + rl];
            carrySp += double(sp.x) - double(sp.y);
            carryHu += hu;
        }
    }
    // per step: the minimum WEPL over this group's rays that leave material (fo > 0), from the
    // 16 lanes of each row (one thread per ray for the step's segment; 16 rays per work-group:
    // 576 work-groups for the head beam spread over the 256 CUs more evenly than 288 of 32 rays)
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        const int sg = s0 + c;
#pragma unroll
        for (int j = 0; j < kTraceSeg; ++j) {
            const int k = sg * kTraceSeg + j;
            // uniform over the row: its 16 lanes share t
            if (sg >= s1 || k >= g.K) continue;
            const float m = row_min((rayIn && fov > 0) ? wv[c * kTraceSeg + j] + 0.0f : 1.0e20f);  // -0 -> +0
            if (rl == 0) partMin[size_t(k) * partPitch + blockIdx.x] = m;
        }
    }
    // the last work-group also fills the pitch's pad columns
    const int groups = (nray + kTraceRays - 1) / kTraceRays;
    if (int(blockIdx.x) == groups - 1)
        for (int i = threadIdx.x; i < g.K * (partPitch - groups); i += blockDim.x)
            partMin[size_t(i / (partPitch - groups)) * partPitch + groups + i % (partPitch - groups)] = 1.0e20f;
    __syncthreads();
    if (t == 0) {
        int fiv = INT_MAX;
        if (rayIn) {
            fiv = max(s_cnt[rl], 1);
            fi[ray] = fiv;
            fo[ray] = fov;
        }
        int mn = fiv, mx = fov;
#pragma unroll
        for (int o = kTraceRays / 2; o > 0; o >>= 1) {
            mn = min(mn, __shfl_xor(mn, o, kTraceRays));
            mx = max(mx, __shfl_xor(mx, o, kTraceRays));
        }
        if (rl == 0) {
            partFi[blockIdx.x] = mn;
            partFo[blockIdx.x] = mx;
        }
    }
}
template <int CH>
struct TraceK {
    static constexpr int NT = kTraceRays * kTraceThreadsPerRay;
    struct Args {
        FanGeom g;
        const float *ct;
        LutDev lut;
        float *density;
        float *wepl;
        int *fi;
        int *fo;
        float *partMin;  // [K][partPitch]
        int partPitch;
        int *partFi;     // [work-groups]
        int *partFo;
    };
    static __device__ __forceinline__ void run(const Args &a, const uint3 b, const uint3 d) {
        (void)d;
        trace_body<CH>(b, a.g, a.ct, a.lut, a.density, a.wepl, a.fi, a.fo, a.partMin, a.partPitch, a.partFi, a.partFo);
    }
};


// WEQ-volume input in place of the trace (fillBevFromWeqVolumeKernel, wrapper:1923-1988), thread
// per ray: the ray's WEQ column is the one nearest its gantry (x, y) (rounded half away from
// zero); WEPL(k) = the column's sample k (the last one read past the column or off the lattice),
// density = max(0, WEPL(k) - WEPL(k-1)) / step; entry / exit steps from the first / last step
// with density > 1e-4 (+1; none: K and 0).
__global__ __launch_bounds__(256) void k_weq_fill(FanGeom g, WeqArgs a, float *__restrict__ density,
                                                  float *__restrict__ wepl, int *__restrict__ fi,
                                                  int *__restrict__ fo) {
    const int r = int(blockIdx.x) * 256 + int(threadIdx.x);
    const int plane = g.rx * g.ry;
    if (r >= plane) return;
    const int x = r % g.rx, y = r / g.rx;
    const float rayX = a.rayX0 + float(x) * a.rayDx, rayY = a.rayY0 + float(y) * a.rayDy;
    const float qx = fabsf(a.dx) > 0.0f ? roundf((rayX - a.x0) / a.dx) : -1.0f;
    const float qy = fabsf(a.dy) > 0.0f ? roundf((rayY - a.y0) / a.dy) : -1.0f;
    const bool in = qx >= 0.0f && qx < float(a.nx) && qy >= 0.0f && qy < float(a.ny);
    const float *col = in ? a.body + (size_t(int(qx)) * a.ny + size_t(int(qy))) * a.nSteps : nullptr;
    const float step = a.stepMm > 0.0f ? a.stepMm : 1.0f;
    float prev = 0.0f, last = 0.0f;
    int first = -1, lastMat = -1;
    for (int k = 0; k < g.K; ++k) {
        float w = last;
        if (in && k < a.active) {
            w = col[k];
            last = w;
        }
        const float d = fmaxf(0.0f, w - prev) / step;
        wepl[size_t(k) * plane + r] = w;
        density[size_t(k) * plane + r] = d;
        if (d > 1.0e-4f) {
            if (first < 0) first = k;
            lastMat = k;
        }
        prev = w;
    }
    fi[r] = first >= 0 ? first + 1 : g.K;
    fo[r] = lastMat >= 0 ? lastMat + 1 : 0;
}

// blocks [0, K): min WEPL over rays that leave material; block K: entry/exit step extremes.
__global__ __launch_bounds__(256) void k_trace_reduce(FanGeom g, const float *__restrict__ wepl,
                                                      const int *__restrict__ fi, const int *__restrict__ fo,
                                                      float *__restrict__ weplMin, BeamScalars *__restrict__ bs) {
    __shared__ float s_f[4];
    __shared__ int s_a[4], s_b[4];
    const int n = g.rx * g.ry;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (blockIdx.x < unsigned(g.K)) {
        const size_t base = size_t(blockIdx.x) * n;
        float m = 1.0e20f;  // Macro.cuh INF as float
        for (int r = threadIdx.x; r < n; r += blockDim.x)
            if (fo[r] > 0) m = fminf(m, wepl[base + r]);
        m = wave_min(m);
        if (lane == 0) s_f[w] = m;
        __syncthreads();
        if (threadIdx.x == 0) weplMin[blockIdx.x] = fminf(fminf(s_f[0], s_f[1]), fminf(s_f[2], s_f[3]));
    } else {
        int mn = INT_MAX, mx = INT_MIN;
        for (int r = threadIdx.x; r < n; r += blockDim.x) { mn = min(mn, fi[r]); mx = max(mx, fo[r]); }
        mn = wave_min_i(mn);
        mx = wave_max_i(mx);
        if (lane == 0) { s_a[w] = mn; s_b[w] = mx; }
        __syncthreads();
        if (threadIdx.x == 0) {
            bs->beamFirstInside = min(min(s_a[0], s_a[1]), min(s_a[2], s_a[3]));
            bs->beamFirstOutside = max(max(s_b[0], s_b[1]), max(s_b[2], s_b[3]));
        }
    }
}

// Distal cut-offs, entry plane and entry sigmas (host formulas, evaluated without
// contraction so the integers match the host's float arithmetic).
__device__ __forceinline__ void beam_setup_body(const uint3 blockIdx, const uint3 gridDim, FanGeom g, const float *__restrict__ weplMin,
                                                    const LayerHost *__restrict__ lh, BeamScalars *__restrict__ bs,
                                                    LayerScalars *__restrict__ ls, const float *__restrict__ subspots,
                                                    int maxSub, float *__restrict__ subEntry, float nucFactor,
                                                    int *__restrict__ classCount, int *__restrict__ hist) {
#pragma clang fp contract(off)
    const int inside = bs->beamFirstInside, outside = bs->beamFirstOutside;
    const int fpa = first_larger_ordered(weplMin, g.K, g.maxCutoffWepl);
    const int gp = min(fpa, outside);
    const float entryZ = float(inside) * g.fd[2] + g.fc[2];
    const float pxX = 1.0f - entryZ / g.dist[0], pxY = 1.0f - entryZ / g.dist[1];
    if (threadIdx.x == 0) {
        bs->firstPastCutoffAll = fpa;
        bs->guaranteedPassive = gp;
        bs->entryZ = entryZ;
        bs->pxMultX = pxX;
        bs->pxMultY = pxY;
    }
    if (threadIdx.x < 16) classCount[threadIdx.x] = 0;
    for (int i = threadIdx.x; i < g.numLayers * (kMaxRadDev + 2); i += blockDim.x) hist[i] = 0;
    for (int l = threadIdx.x; l < g.numLayers; l += blockDim.x) {
        const LayerHost L = lh[l];
        const int local = first_larger_ordered(weplMin, g.K, L.cutoffWepl);
        LayerScalars s;
        s.afterLast = min(local, gp);
        s.passiveRaw = INT_MIN;
        s.passive = 0;
        s.mainEnd = 0;
        s.group = -1;
        const float base = L.convA * entryZ * entryZ + L.convB * entryZ;
        const float vx = base + L.sigmaX * L.sigmaX, vy = base + L.sigmaY * L.sigmaY;
        s.entrySigX = sqrtf(fmaxf(0.0f, vx)) * nucFactor;
        s.entrySigY = sqrtf(fmaxf(0.0f, vy)) * nucFactor;
        s.pad0 = 0;
        ls[l] = s;
        for (int q = 0; q < maxSub && subspots; ++q) {
            const float *ss = subspots + (size_t(l) * maxSub + q) * 5;
            const float sxq = ss[3], syq = ss[4];
            subEntry[(size_t(l) * maxSub + q) * 2 + 0] = sqrtf(fmaxf(0.0f, base + sxq * sxq)) * nucFactor;
            subEntry[(size_t(l) * maxSub + q) * 2 + 1] = sqrtf(fmaxf(0.0f, base + syq * syq)) * nucFactor;
        }
    }
}
struct BeamSetupK {
    static constexpr int NT = 256;
    struct Args {
        FanGeom g;
        const float *weplMin;
        const LayerHost *lh;
        BeamScalars *bs;
        LayerScalars *ls;
        const float *subspots;
        int maxSub;
        float *subEntry;
        float nucFactor;
        int *classCount;
        int *hist;
    };
    static __device__ __forceinline__ void run(const Args &a, const uint3 b, const uint3 d) { beam_setup_body(b, d, a.g, a.weplMin, a.lh, a.bs, a.ls, a.subspots, a.maxSub, a.subEntry, a.nucFactor, a.classCount, a.hist); }
};


// ---------------------------------------------------------------------------------------
// spot lattice -> ray weights
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ float erf_gather(const float *__restrict__ in, int inN, int inStride, int o, float sigma,
                                            float pixelSp, float off, float delta) {
    const float sEff = sigma / pixelSp;
    const float rS = rsqrtf(2.0f) / sEff;
    const float reach = 3.0f * sEff + 0.5f;
    int ci = int(ceilf((float(o) - reach - off) / delta));
    float dist = float(ci) * delta + off - float(o);
    float res = 0.0f;
    while (dist < reach) {
        if (ci >= 0 && ci < inN) res += 0.5f * (erff((dist + 0.5f) * rS) - erff((dist - 0.5f) * rS)) * in[size_t(ci) * inStride];
        ++ci;
        dist = float(ci) * delta + off - float(o);
    }
    return res;
}

// grid: (ceil(rx/64), spotY, L * nSub); interm[(l*nSub + s)][spotY][rx]
__device__ __forceinline__ void conv_x_body(const uint3 blockIdx, const uint3 gridDim, ConvArgs a, const float *__restrict__ spot, float *__restrict__ interm,
                                               const LayerScalars *__restrict__ ls, const BeamScalars *__restrict__ bs,
                                               const float *__restrict__ subspots, const float *__restrict__ subEntry) {
    const int o = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y;
    const int l = blockIdx.z / a.nSub, s = blockIdx.z % a.nSub;
    if (o >= a.rx) return;
    float sigma = ls[l].entrySigX, off = a.offX, scale = 1.0f;
    if (a.subspotMode) {
        const float *ss = subspots + (size_t(l) * a.nSub + s) * 5;
        scale = ss[2];
        sigma = subEntry[(size_t(l) * a.nSub + s) * 2];
        // subspot offsets shift the lattice origin (wrapper:3802-3806)
        off = (a.spotOffX + ss[0] / a.lenToMm - a.cpbCornerX) / a.cpbResX;
        if (!(scale > 1.0e-8f) || !(ss[3] > 1.0e-8f) || !(ss[4] > 1.0e-8f)) scale = 0.0f;
    }
    float r = 0.0f;
    if (scale != 0.0f) {
        const float *row = spot + (size_t(l) * a.spotY + y) * a.spotX;
        r = scale * erf_gather(row, a.spotX, 1, o, sigma, a.cpbResX * bs->pxMultX, off, a.deltaX);
    }
    interm[((size_t(blockIdx.z) * a.spotY + y) * a.rx) + o] = r;
}
struct ConvXK {
    static constexpr int NT = 64;
    struct Args {
        ConvArgs a;
        const float *spot;
        float *interm;
        const LayerScalars *ls;
        const BeamScalars *bs;
        const float *subspots;
        const float *subEntry;
    };
    static __device__ __forceinline__ void run(const Args &a, const uint3 b, const uint3 d) { conv_x_body(b, d, a.a, a.spot, a.interm, a.ls, a.bs, a.subspots, a.subEntry); }
};


// grid: (ceil(rx/64), ry, L); sums the subspot planes in subspot order
__device__ __forceinline__ void conv_y_body(const uint3 blockIdx, const uint3 gridDim, ConvArgs a, const float *__restrict__ interm, float *__restrict__ rayW,
                                               const LayerScalars *__restrict__ ls, const BeamScalars *__restrict__ bs,
                                               const float *__restrict__ subspots, const float *__restrict__ subEntry) {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int o = blockIdx.y, l = blockIdx.z;
    if (x >= a.rx) return;
    float acc = 0.0f;
    for (int s = 0; s < a.nSub; ++s) {
        float sigma = ls[l].entrySigY, off = a.offY;
        if (a.subspotMode) {
            const float *ss = subspots + (size_t(l) * a.nSub + s) * 5;
            if (!(ss[2] > 1.0e-8f) || !(ss[3] > 1.0e-8f) || !(ss[4] > 1.0e-8f)) continue;
            sigma = subEntry[(size_t(l) * a.nSub + s) * 2 + 1];
            off = (a.spotOffY + ss[1] / a.lenToMm - a.cpbCornerY) / a.cpbResY;
        }
        const float *col = interm + (size_t(l) * a.nSub + s) * a.spotY * a.rx + x;
        acc += erf_gather(col, a.spotY, a.rx, o, sigma, a.cpbResY * bs->pxMultY, off, a.deltaY);
    }
    rayW[(size_t(l) * a.ry + o) * a.rx + x] = acc;
}
struct ConvYK {
    static constexpr int NT = 64;
    struct Args {
        ConvArgs a;
        const float *interm;
        float *rayW;
        const LayerScalars *ls;
        const BeamScalars *bs;
        const float *subspots;
        const float *subEntry;
    };
    static __device__ __forceinline__ void run(const Args &a, const uint3 b, const uint3 d) { conv_y_body(b, d, a.a, a.interm, a.rayW, a.ls, a.bs, a.subspots, a.subEntry); }
};


// CPB weights as a gather: each cell sums the subspots whose 3-sigma cell window covers it,
// in subspot order. Window membership is the reference's floor/ceil rule.
__global__ __launch_bounds__(256) void k_subspot_cpb(CpbArgs a, const float *__restrict__ sub, float *__restrict__ cpb) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y;
    if (cell >= a.nx * a.ny) return;
    const int cx = cell % a.nx, cy = cell / a.nx;
    float acc = 0.0f;
    for (int s = 0; s < a.nSub; ++s) {
        const float *q = sub + (size_t(l) * a.nSub + s) * 5;
        const float w = q[2], sX = q[3], sY = q[4];
        // SubspotInfo position (common.cuh:186-194)
        const float px = a.src[0] + a.bd[0] * a.sad + a.bx[0] * q[0] + a.by[0] * q[1];
        const float py = a.src[1] + a.bd[1] * a.sad + a.bx[1] * q[0] + a.by[1] * q[1];
        const float pz = a.src[2] + a.bd[2] * a.sad + a.bx[2] * q[0] + a.by[2] * q[1];
        const float dxs = px - a.src[0], dys = py - a.src[1], dzs = pz - a.src[2];
        if (!(sqrtf(dxs * dxs + dys * dys + dzs * dzs) > 1e-6f)) continue;
        if (!(w > 1e-6f && sX > 1e-6f && sY > 1e-6f)) continue;
        const int x0 = max(0, int(floorf((px - 3.0f * sX - a.cx) / a.rx_)));
        const int x1 = min(a.nx - 1, int(ceilf((px + 3.0f * sX - a.cx) / a.rx_)));
        const int y0 = max(0, int(floorf((py - 3.0f * sY - a.cy) / a.ry_)));
        const int y1 = min(a.ny - 1, int(ceilf((py + 3.0f * sY - a.cy) / a.ry_)));
        if (cx < x0 || cx > x1 || cy < y0 || cy > y1) continue;
        const float ccx = a.cx + (float(cx) + 0.5f) * a.rx_, ccy = a.cy + (float(cy) + 0.5f) * a.ry_;
        const float ddx = ccx - px, ddy = ccy - py;
        const float r2x = 1.0f / (1.41421356f * sX), r2y = 1.0f / (1.41421356f * sY);
        const float ex = erff((ddx + 0.5f * a.rx_) * r2x) - erff((ddx - 0.5f * a.rx_) * r2x);
        const float ey = erff((ddy + 0.5f * a.ry_) * r2y) - erff((ddy - 0.5f * a.ry_) * r2y);
        acc += 0.25f * ex * ey * w;
    }
    cpb[(size_t(l) * a.ny + cy) * a.nx + cx] = acc;
}

__global__ __launch_bounds__(256) void k_cpb_to_rays(CpbArgs a, const float *__restrict__ cpb, float *__restrict__ rays,
                                                     int rnx, int rny, float rcx, float rcy, float rrx, float rry) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y;
    if (r >= rnx * rny) return;
    const int x = r % rnx, y = r / rnx;
    const float x0 = rcx + float(x) * rrx, x1 = x0 + rrx, y0 = rcy + float(y) * rry, y1 = y0 + rry;
    int cx0 = int(floorf((x0 - a.cx) / a.rx_)), cx1 = int(ceilf((x1 - a.cx) / a.rx_)) - 1;
    int cy0 = int(floorf((y0 - a.cy) / a.ry_)), cy1 = int(ceilf((y1 - a.cy) / a.ry_)) - 1;
    cx0 = max(cx0, 0); cy0 = max(cy0, 0);
    cx1 = min(cx1, a.nx - 1); cy1 = min(cy1, a.ny - 1);
    const float area = a.rx_ * a.ry_;
    float acc = 0.0f;
    if (area > 0.0f)
        for (int cy = cy0; cy <= cy1; ++cy) {
            const float c0 = a.cy + float(cy) * a.ry_, c1 = c0 + a.ry_;
            const float oy = fmaxf(0.0f, fminf(y1, c1) - fmaxf(y0, c0));
            if (oy <= 0.0f) continue;
            for (int cx = cx0; cx <= cx1; ++cx) {
                const float d0 = a.cx + float(cx) * a.rx_, d1 = d0 + a.rx_;
                const float ox = fmaxf(0.0f, fminf(x1, d1) - fmaxf(x0, d0));
                if (ox <= 0.0f) continue;
                acc += cpb[(size_t(l) * a.ny + cy) * a.nx + cx] * (ox * oy) / area;
            }
        }
    rays[(size_t(l) * rny + y) * rnx + x] = acc;
}

// Tile radius of a 32x8 tile from its smallest rSigmaEff (complete_superposition.cu:242):
// min(int(KS_SIGMA_CUTOFF / (sqrt(2) min) + 1/2), MAX_RAD + 1); NaN minimum -> 0.
__device__ __forceinline__ int tile_radius(float m) {
    if (m != m) return 0;
    const float q = 3.0f / (1.41421356237f * m) + 0.5f;  // KS_SIGMA_CUTOFF / (sqrt(2) min) + 1/2
    const int r = int(q);
    return r < kMaxRadDev + 1 ? r : kMaxRadDev + 1;
}

// ---------------------------------------------------------------------------------------
// IDD and sigma
// ---------------------------------------------------------------------------------------
// Hardware transcendentals (about 1 ulp): the fill's step is VALU-issue bound, and the
// correctly rounded library forms cost 10-80 instructions each.
__device__ __forceinline__ float rcp_hw(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float sqrt_hw(float x) { return __builtin_amdgcn_sqrtf(x); }
// x^e for x > 0 as 2^(e log2 x)
__device__ __forceinline__ float pow_hw(float x, float e) { return __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(x)); }

// Bilinear lookup in a table row pair staged in LDS: the same arithmetic as lut2d_clamp for
// a fixed row coordinate (rows r0 = row y0, r1 = row y1, fraction ay).
// The same with the two rows interleaved, (r0[i], r1[i]) one float2: two 8-byte reads.
__device__ __forceinline__ float lut_rows2(const float2 *__restrict__ r, int nx, float ay, float xb) {
    const float fx = floorf(xb);
    const float ax = xb - fx;
    int x0 = int(fx);
    const int x1 = min(max(x0 + 1, 0), nx - 1);
    x0 = min(max(x0, 0), nx - 1);
    const float2 a = r[x0], b = r[x1];
    const float v0 = (1.0f - ax) * a.x + ax * b.x;
    const float v1 = (1.0f - ax) * a.y + ax * b.y;
    return (1.0f - ay) * v0 + ay * v1;
}
__device__ __forceinline__ float lut_rows(const float *__restrict__ r0, const float *__restrict__ r1, int nx,
                                          float ay, float xb) {
    const float fx = floorf(xb);
    const float ax = xb - fx;
    int x0 = int(fx);
    const int x1 = min(max(x0 + 1, 0), nx - 1);
    x0 = min(max(x0, 0), nx - 1);
    const float v0 = (1.0f - ax) * r0[x0] + ax * r0[x1];
    const float v1 = (1.0f - ax) * r1[x0] + ax * r1[x1];
    return (1.0f - ay) * v0 + ay * v1;
}

// One block per (32x8 tile, layer); thread per ray of the tile. Besides IDD / rSigmaEff per
// step it reduces each step's tile: code = radius | live << 8 (the tileRadCalc of
// complete_superposition.cu:197-255, fused so the field is not read back).
//
// The recurrence (idd_sigma.cu:130-230) carries only sigma^2 and its three increments from step
// to step; WEPL, density, the cumulative dose and the stopping power are functions of the loaded
// WEPL / density. So each step is written branch-free (selects on the ray's live flag), WEPL and
// density are loaded a chunk of kFillChunk steps ahead, and the tile codes of a chunk are
// written after one barrier per chunk: the compiler can then overlap the independent parts of
// the steps of a chunk, which is what bounds the longest layer's blocks. Eight steps: 82 -> 80 us
// on the head beam against four (fewer barriers); 12 and 16 were slower again.
constexpr int kFillChunk = 8;

template <bool HALO>
__device__ __forceinline__ void fill_body(const uint3 blockIdx, const uint3 gridDim, FanGeom g, LutDev lut, const LayerHost *__restrict__ lh,
                                              LayerScalars *__restrict__ ls, const BeamScalars *__restrict__ bs,
                                              const float *__restrict__ density, const float *__restrict__ wepl,
                                              const float *__restrict__ rayW, const int *__restrict__ fi,
                                              const int *__restrict__ fo, float2 *__restrict__ iddSigma,
                                              int *__restrict__ firstPassive, FillConsts fc, HaloFill hf,
                                              int *__restrict__ tileCode) {
    // LDS: 1/X0 LUT, per-step beam geometry (ray spacing factor, voxel mass factor), the layer's
    // two CIDD rows, then (halo) its nuclear table rows
    extern __shared__ float s_lds[];
    __shared__ float s_min[2][kFillChunk][4];
    __shared__ int s_live[2][kFillChunk][4];
    const int ntx = g.sx / 32, nt = ntx * (g.sy / 8);
    const int t = blockIdx.x, l = blockIdx.y;
    const float eIdx = lh[l].energyIdx;
    const float eFl = floorf(eIdx);
    const float ay = eIdx - eFl;
    const int ey0 = min(max(int(eFl), 0), lut.nE - 1), ey1 = min(max(int(eFl) + 1, 0), lut.nE - 1);
    const int nS = lut.nS, nRrl = lut.nRrl;
    float *s_rrl = s_lds;
    float *s_vw = s_rrl + nRrl, *s_mp = s_vw + g.K;
    // the layer's two CIDD rows interleaved, 8-byte aligned (one float of padding at most)
    float2 *c01 = reinterpret_cast<float2 *>(s_mp + g.K + ((nRrl + 2 * g.K) & 1));
    float *nw0 = reinterpret_cast<float *>(c01 + nS), *nw1 = nw0 + nS, *ns0 = nw1 + nS, *ns1 = ns0 + nS;
    for (int i = threadIdx.x; i < nRrl; i += blockDim.x) s_rrl[i] = lut.rrl[i];
    for (int i = threadIdx.x; i < nS; i += blockDim.x) {
        c01[i] = make_float2(lut.cidd[size_t(ey0) * nS + i], lut.cidd[size_t(ey1) * nS + i]);
        if (HALO) {
            nw0[i] = hf.nucWeightLut[size_t(ey0) * nS + i];
            nw1[i] = hf.nucWeightLut[size_t(ey1) * nS + i];
            ns0[i] = hf.nucSqSigmaLut[size_t(ey0) * nS + i];
            ns1[i] = hf.nucSqSigmaLut[size_t(ey1) * nS + i];
        }
    }
    for (int k = threadIdx.x; k < g.K; k += blockDim.x) {
        // half the summed fan ray spacings at the step's depth, and the voxel volume at step k
        const float zk = g.fc[2] + float(k) * g.fd[2];
        const float vwx = g.fd[0] * (1.0f - zk / g.dist[0]);
        const float vwy = g.fd[1] * (1.0f - zk / g.dist[1]);
        s_vw[k] = 0.5f * (vwx + vwy);
        const float fk = float(k);
        s_mp[k] = g.volConst + fk * g.volLin + fk * fk * g.volSq;
    }
    const int x = (t % ntx) * 32 + (threadIdx.x & 31), y = (t / ntx) * 8 + (threadIdx.x >> 5);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const bool rayIn = x < g.rx && y < g.ry;
    const int nray = g.rx * g.ry;
    const int ray = rayIn ? y * g.rx + x : 0;
    const LayerHost L = lh[l];
    const unsigned first = unsigned(bs->beamFirstInside);
    const unsigned pAfter = unsigned(ls[l].afterLast);
    const unsigned firstIn = rayIn ? unsigned(fi[ray]) : 0u;
    unsigned afterLast = rayIn ? unsigned(min(fo[ray], int(pAfter))) : 0u;
    const float w0 = rayIn ? (fc.unitWeight ? 1.0f : rayW[size_t(l) * nray + ray]) : 0.0f;
    bool live = rayIn;
    if (w0 < 1e-6f || afterLast < first) { live = false; afterLast = 0; }
    float spOld = 0.f, cdOld = 0.f;
    float incScat = 0.f, incincScat = 0.f;
    float incDiv = L.airLin + (2.0f * float(first) - 1.0f) * L.airQuad;
    float sigmaSq = -incDiv;
    const float stepLength = fabsf(g.fd[2]);
    const float scatStep = 2.0f * stepLength * stepLength;
    const size_t plane = size_t(nray);
    float2 *out = iddSigma + (size_t(l) * g.K * g.sy + y) * g.sx + x;
    const size_t outStep = size_t(g.sx) * g.sy;
    int *codes = tileCode + size_t(l) * g.K * nt + t;
    // halo cell of this ray (the spot it centres) and the halo entry sigma^2 (wrapper:4324-4329)
    const int cell = (HALO && rayIn) ? hf.rayToNuc[ray] : -1;
    float nucRayW = 0.0f, entrySigmaSq = 0.0f;
    float2 *nout = nullptr;
    if (HALO && cell >= 0) {
        nucRayW = hf.weights[size_t(l) * hf.nucPlane + cell];
        nout = hf.out + size_t(l) * g.K * hf.nucPlane + cell;
        {
#pragma clang fp contract(off)
            const float ez = bs->entryZ;
            entrySigmaSq = L.airA * ez * ez + L.airC * ez + L.sigmaX * L.sigmaX;
        }
    }
    __syncthreads();

    // one step of the recurrence for this ray (idd_sigma.cu:130-230); outputs are zero / inf
    // once the ray has stopped and before its first inside step
    auto step = [&](unsigned k, float sp, float rho, float &res, float &rs, float &nres, float &nrs) {
        const float cd = lut_rows2(c01, nS, ay, sp * L.scaleFact);
        const float mid = 0.5f * (sp + spOld);
        const bool pre = sp < L.peakDepth;
        const float resE = fc.rangeCoef * pow_hw(L.peakDepth - mid, fc.rangeExp);
        const float M = fc.restMass;
        const float betaP = resE + M - M * M * rcp_hw(resE + M);
        const float rRl = rho * lut1d_clamp(s_rrl, nRrl, rho * lut.rrlScale);
        const float thetaSq = fc.eRefSq * rcp_hw(betaP * betaP) * stepLength * rRl;
        const float incSum = incScat + incDiv;
        const float iis = incincScat + thetaSq * scatStep;
        const float sgPost = fc.suppressDistalDip ? sigmaSq : sigmaSq - 1.5f * incSum * rho;
        const bool adv = live && pre;
        sigmaSq = live ? (pre ? sigmaSq + incSum : sgPost) : sigmaSq;
        incincScat = adv ? iis : incincScat;
        incScat = adv ? incScat + iis : incScat;
        incDiv = adv ? incDiv + 2.0f * L.airQuad : incDiv;
        const float vw = s_vw[k];
        const float rsv = vw * rcp_hw(1.41421356f * (sqrt_hw(sigmaSq) + fc.sigmaDelta));
        const float mass = rho * s_mp[k];
        const float rMass = rcp_hw(mass);
        const float dd = cd - cdOld;
        const bool heavy = mass > 1e-2f;
        float r, nr = 0.0f, nrsv = PBS_INF;
        if (HALO) {
            // idd_sigma.cu:194-213: the halo takes w_nuc of the step's energy
            const float depthMid = mid * L.scaleFact;
            const float nw = lut_rows(nw0, nw1, nS, ay, depthMid);
            r = (1.0f - nw) * w0 * dd * rMass;
            nr = nw * nucRayW * dd * rMass * hf.rSpotDistSq;
            const float s2 = lut_rows(ns0, ns1, nS, ay, depthMid);
            nrsv = hf.spotDist * vw * rcp_hw(1.41421356f * sqrt_hw(sigmaSq + s2 + entrySigmaSq));
            nr = (heavy && cell >= 0) ? nr : 0.0f;
            nrsv = cell >= 0 ? nrsv : PBS_INF;
        } else {
            r = w0 * dd * rMass;
        }
        r = heavy ? r : 0.0f;
        spOld = sp;
        cdOld = cd;
        const bool stop = sp > L.rangeStop || k == afterLast;
        afterLast = (live && stop) ? k : afterLast;
        live = live && !stop;
        const bool zero = !live || int(k) < int(firstIn) - 1;
        res = zero ? 0.0f : r;
        rs = zero ? PBS_INF : rsv;
        nres = zero ? 0.0f : nr;
        nrs = zero ? PBS_INF : nrsv;
    };
    // store a step and stage its per-wave tile reduction: minimum rSigmaEff (NaN ignored,
    // fminf) and whether any pixel is live
    auto emit = [&](unsigned k, int u, int cb, float res, float rs, float nres, float nrs) {
        out[size_t(k) * outStep] = make_float2(res, rs);
        if (HALO && cell >= 0) nout[size_t(k) * hf.nucPlane] = make_float2(nres, nrs);
        const float m = wave_min_uniform(rs);  // every lane is active here
        const bool any = __any(res > 0.0f);
        if (lane == 0) {
            s_min[cb][u][w] = m;
            s_live[cb][u][w] = any ? 1 : 0;
        }
    };
    // tile codes of the n steps of a chunk from the four waves' reductions
    auto flush = [&](unsigned kc, int n, int cb) {
        __syncthreads();
        if (int(threadIdx.x) < n) {
            const int u = threadIdx.x;
            const float mm = fminf(fminf(s_min[cb][u][0], s_min[cb][u][1]), fminf(s_min[cb][u][2], s_min[cb][u][3]));
            const int lv = s_live[cb][u][0] | s_live[cb][u][1] | s_live[cb][u][2] | s_live[cb][u][3];
            codes[size_t(kc + u) * nt] = tile_radius(mm) | (lv ? 256 : 0);
        }
    };

    // WEPL and density of the next chunk are loaded while the current one is computed; loads
    // past the layer's last step are clamped to it (their values are never used)
    const unsigned kLast = pAfter > 0 ? pAfter - 1 : 0;
    float spBuf[kFillChunk], rhoBuf[kFillChunk];
#pragma unroll
    for (int u = 0; u < kFillChunk; ++u) {
        const size_t k = min(first + unsigned(u), kLast);
        spBuf[u] = wepl[k * plane + ray];
        rhoBuf[u] = density[k * plane + ray];
    }
    unsigned kc = first;
    int cb = 0;
    for (; kc + kFillChunk <= pAfter; kc += kFillChunk, cb ^= 1) {
        float spNx[kFillChunk], rhoNx[kFillChunk];
#pragma unroll
        for (int u = 0; u < kFillChunk; ++u) {
            const size_t k = min(kc + kFillChunk + unsigned(u), kLast);
            spNx[u] = wepl[k * plane + ray];
            rhoNx[u] = density[k * plane + ray];
        }
        if (__any(live)) {
#pragma unroll
            for (int u = 0; u < kFillChunk; ++u) {
                float res, rs, nres, nrs;
                step(kc + u, spBuf[u], rhoBuf[u], res, rs, nres, nrs);
                emit(kc + u, u, cb, res, rs, nres, nrs);
            }
        } else {
            // every ray of the wave has stopped: zero outputs, no live pixel
#pragma unroll
            for (int u = 0; u < kFillChunk; ++u) {
                out[size_t(kc + u) * outStep] = make_float2(0.0f, PBS_INF);
                if (HALO && cell >= 0) nout[size_t(kc + u) * hf.nucPlane] = make_float2(0.0f, PBS_INF);
                if (lane == 0) {
                    s_min[cb][u][w] = PBS_INF;
                    s_live[cb][u][w] = 0;
                }
            }
        }
        flush(kc, kFillChunk, cb);
#pragma unroll
        for (int u = 0; u < kFillChunk; ++u) {
            spBuf[u] = spNx[u];
            rhoBuf[u] = rhoNx[u];
        }
    }
    const int rem = int(pAfter) - int(kc);
    if (rem > 0) {
#pragma unroll
        for (int u = 0; u < kFillChunk; ++u) {
            if (u < rem) {
                float res, rs, nres, nrs;
                step(kc + u, spBuf[u], rhoBuf[u], res, rs, nres, nrs);
                emit(kc + u, u, cb, res, rs, nres, nrs);
            }
        }
        flush(kc, rem, cb);
    }
    int passive = INT_MIN;
    if (rayIn) {
        firstPassive[size_t(l) * nray + ray] = int(afterLast);
        passive = int(afterLast);
    }
    // layer passive step = max over rays (wrapper:4950-4979): one atomic per wave
    const int mx = wave_max_i(passive);
    if (lane == 0) atomicMax(&ls[l].passiveRaw, mx);
}
template <bool HALO>
struct FillK {
    static constexpr int NT = 256;
    struct Args {
        FanGeom g;
        LutDev lut;
        const LayerHost *lh;
        LayerScalars *ls;
        const BeamScalars *bs;
        const float *density;
        const float *wepl;
        const float *rayW;
        const int *fi;
        const int *fo;
        float2 *iddSigma;
        int *firstPassive;
        FillConsts fc;
        HaloFill hf;
        int *tileCode;
    };
    static __device__ __forceinline__ void run(const Args &a, const uint3 b, const uint3 d) { fill_body<HALO>(b, d, a.g, a.lut, a.lh, a.ls, a.bs, a.density, a.wepl, a.rayW, a.fi, a.fo, a.iddSigma, a.firstPassive, a.fc, a.hf, a.tileCode); }
};


// Weight rescale of a resident beam (the served form's weight-only update): the fill's
// recurrence does not depend on the ray weight, which only scales the IDD and decides the live
// mask (w >= 1e-6, idd_sigma.cu:78). From the unit-weight field it writes, per ray and step,
// IDD = W x unit IDD and rSigmaEff (inf on dead rays), firstPassive, the tile codes and the
// layer's passive step, as the fill does for weights W. One block per (32x8 tile, layer).
__global__ __launch_bounds__(256) void k_rescale(FanGeom g, LayerScalars *__restrict__ ls,
                                                 const BeamScalars *__restrict__ bs, const float *__restrict__ rayW,
                                                 const float2 *__restrict__ unitIdd, const int *__restrict__ unitFp,
                                                 float2 *__restrict__ iddSigma, int *__restrict__ firstPassive,
                                                 int *__restrict__ tileCode) {
    __shared__ float s_min[2][kFillChunk][4];
    __shared__ int s_live[2][kFillChunk][4];
    const int ntx = g.sx / 32, nt = ntx * (g.sy / 8);
    const int t = blockIdx.x, l = blockIdx.y;
    const int x = (t % ntx) * 32 + (threadIdx.x & 31), y = (t / ntx) * 8 + (threadIdx.x >> 5);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const bool rayIn = x < g.rx && y < g.ry;
    const int nray = g.rx * g.ry;
    const int ray = rayIn ? y * g.rx + x : 0;
    const unsigned first = unsigned(bs->beamFirstInside);
    const unsigned pAfter = unsigned(ls[l].afterLast);
    const float W = rayIn ? rayW[size_t(l) * nray + ray] : 0.0f;
    const bool live = rayIn && W >= 1e-6f;
    const size_t outStep = size_t(g.sx) * g.sy;
    const size_t base = (size_t(l) * g.K * g.sy + y) * g.sx + x;
    int *codes = tileCode + size_t(l) * g.K * nt + t;
    for (unsigned kc = first, cb = 0; kc < pAfter; kc += kFillChunk, cb ^= 1u) {
        const int n = min(int(pAfter - kc), kFillChunk);
#pragma unroll
        for (int u = 0; u < kFillChunk; ++u) {
            if (u < n) {
                const size_t i = base + size_t(kc + u) * outStep;
                const float2 v = unitIdd[i];
                const float res = live ? W * v.x : 0.0f;
                const float rs = live ? v.y : PBS_INF;
                iddSigma[i] = make_float2(res, rs);
                const float m = wave_min_uniform(rs);  // every lane is active here
                const bool any = __any(res > 0.0f);
                if (lane == 0) {
                    s_min[cb][u][w] = m;
                    s_live[cb][u][w] = any ? 1 : 0;
                }
            }
        }
        __syncthreads();
        if (int(threadIdx.x) < n) {
            const int u = threadIdx.x;
            const float mm = fminf(fminf(s_min[cb][u][0], s_min[cb][u][1]), fminf(s_min[cb][u][2], s_min[cb][u][3]));
            const int lv = s_live[cb][u][0] | s_live[cb][u][1] | s_live[cb][u][2] | s_live[cb][u][3];
            codes[size_t(kc + u) * nt] = tile_radius(mm) | (lv ? 256 : 0);
        }
    }
    int passive = INT_MIN;
    if (rayIn) {
        const int fp = live ? unitFp[size_t(l) * nray + ray] : 0;
        firstPassive[size_t(l) * nray + ray] = fp;
        passive = fp;
    }
    const int mx = wave_max_i(passive);
    if (lane == 0) atomicMax(&ls[l].passiveRaw, mx);
}

// A layer's passive step from its first passive step (wrapper:4984-4996).
__device__ __forceinline__ int layer_passive(const FanGeom &g, const LayerScalars &L, int inside) {
    int pas = L.passiveRaw;
    if (pas < inside + 1) pas = inside + 1;
    if (pas < 0) pas = inside + 1;
    if (pas > L.afterLast) pas = L.afterLast;
    if (pas > g.K) pas = g.K;
    return pas;
}

// Radius histogram of one layer from its tile codes over the superposed steps
// [beamFirstInside, passive) (complete_superposition.cu:197-255, 400-463); one block per layer.
// Layer blockIdx.x's radius histogram of its tile codes over [beamFirstInside, passive) into
// s_h (complete at return: the caller's threads may read all of it).
__device__ __forceinline__ void tile_hist_into(const uint3 blockIdx, const FanGeom &g, const int *__restrict__ tileCode,
                                               const LayerScalars *__restrict__ ls, const BeamScalars *__restrict__ bs,
                                               int *s_h) {
    const int l = blockIdx.x;
    const int nt = (g.sx / 32) * (g.sy / 8);
    if (threadIdx.x < kMaxRadDev + 2) s_h[threadIdx.x] = 0;
    __syncthreads();
    const int k0 = bs->beamFirstInside, k1 = layer_passive(g, ls[l], k0);  // = ls[l].passive
    const int n = k1 > k0 ? (k1 - k0) * nt : 0;
    const int *codes = tileCode + (size_t(l) * g.K + k0) * nt;
    for (int i = threadIdx.x; i < n; i += blockDim.x) atomicAdd(&s_h[codes[i] & 255], 1);
    __syncthreads();
}


// Transfer box of one passive step (wrapper:5343-5395): 8 fan corners -> dose index in float,
// floor/ceil, x start rounded down to 32; the launch covers whole 32x8 thread blocks from the
// start, so the box end is rounded up to those blocks and clipped to the grid.
__device__ void transfer_box(const FanGeom &g, int inside, int passive, int st[3], int en[3], int mi[3]) {
#pragma clang fp contract(off)
    const float xv[2] = {-32.0f, float(g.sx + 32 - 1)};
    const float yv[2] = {-32.0f, float(g.sy + 32 - 1)};
    const float zv[2] = {float(inside), float(passive - 1)};
    float mx[3] = {-1.f, -1.f, -1.f}, mn[3] = {100000.f, 100000.f, 100000.f};
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c) {
                float f0 = xv[c] * g.fd[0] + g.fc[0], f1 = yv[b] * g.fd[1] + g.fc[1];
                const float f2 = zv[a] * g.fd[2] + g.fc[2];
                f0 *= (1.0f - f2 / g.dist[0]);
                f1 *= (1.0f - f2 / g.dist[1]);
                for (int r = 0; r < 3; ++r) {
                    const float q = (g.g2dose[3 * r] * f0 + g.g2dose[3 * r + 1] * f1 + g.g2dose[3 * r + 2] * f2) + g.g2doseOff[r];
                    mx[r] = fmaxf(mx[r], q);
                    mn[r] = fminf(mn[r], q);
                }
            }
    st[0] = max((int(floorf(mn[0])) / 32) * 32, 0);
    st[1] = max(int(floorf(mn[1])), 0);
    st[2] = max(int(floorf(mn[2])), 0);
    for (int r = 0; r < 3; ++r) mi[r] = min(int(ceilf(mx[r])), g.doseDims[r] - 1);
    en[0] = min(st[0] + ((mi[0] - st[0] + 1 + 31) / 32) * 32, g.doseDims[0]);
    en[1] = min(st[1] + ((mi[1] - st[1] + 1 + 7) / 8) * 8, g.doseDims[1]);
    en[2] = mi[2] + 1;
    for (int r = 0; r < 3; ++r) if (en[r] < st[r]) en[r] = st[r];
}

// BEV plane sampled by dose voxel (x, y, z): the fan step minus beamFirstInside.
__device__ __forceinline__ float sampled_plane(const FanGeom &g, int inside, float x, float y, float z) {
    const float gz = g.d2g[6] * x + g.d2g[7] * y + g.d2g[8] * z + g.d2gOff[2];
    return (gz - g.fc[2]) / g.fd[2] - float(inside);
}

// Layer passive steps, layer groups, group boxes and tail planes (one block of 256 threads;
// the groups are built in LDS and written out once).
__device__ __forceinline__ void layer_finalize_body(const uint3 blockIdx, const uint3 gridDim, FanGeom g, LayerScalars *__restrict__ ls,
                                                        GroupScalars *__restrict__ gs, BeamScalars *__restrict__ bs,
                                                        int *__restrict__ tailPlaneGroup, int *__restrict__ tailPlaneK,
                                                        int tailCap, int *__restrict__ planeRM) {
    for (int i = threadIdx.x; i < g.K + tailCap; i += blockDim.x) planeRM[i] = -1;  // per-plane max layout radius
    __shared__ int s_pas[kMaxLayersDev];
    __shared__ int s_first[kMaxLayersDev];
    __shared__ int s_group[kMaxLayersDev];
    __shared__ GroupScalars s_gs[kMaxLayersDev];
    __shared__ int s_m, s_total;
    const int inside = bs->beamFirstInside;
    const int L = g.numLayers;
    for (int l = threadIdx.x; l < L; l += blockDim.x) {
        const int pas = layer_passive(g, ls[l], inside);
        ls[l].passive = pas;
        s_pas[l] = pas > inside ? pas : -1;  // layers with no superposed step take no part
    }
    __syncthreads();
    // groups = distinct passive steps in ascending order; group of a layer = rank of its step
    for (int l = threadIdx.x; l < L; l += blockDim.x) {
        int first = s_pas[l] >= 0 ? 1 : 0;
        for (int q = 0; q < l && first; ++q) first = (s_pas[q] != s_pas[l]);
        s_first[l] = first;
    }
    __syncthreads();
    for (int l = threadIdx.x; l < L; l += blockDim.x) {
        int rank = -1;
        if (s_pas[l] >= 0) {
            rank = 0;
            for (int q = 0; q < L; ++q) rank += (s_first[q] && s_pas[q] < s_pas[l]) ? 1 : 0;
            if (s_first[l]) s_gs[rank].passive = s_pas[l];
        }
        s_group[l] = rank;
    }
    if (threadIdx.x == 0) {
        int m = 0;
        for (int q = 0; q < L; ++q) m += s_first[q];
        s_m = m;
    }
    __syncthreads();
    const int m = s_m;
    for (int j = threadIdx.x; j < m; j += blockDim.x) {
        GroupScalars G = s_gs[j];
        transfer_box(g, inside, G.passive, G.boxStart, G.boxEnd, G.boxMax);
        // main BEV keeps steps below passive - inside - 1: no voxel outside the group's box
        // samples them; the tail covers the rest of what voxels inside the box can sample.
        G.tailStart = max(inside, G.passive - inside - 1);
        float tmax = -1.0e30f;
        for (int c = 0; c < 8; ++c) {
            const float x = float((c & 1) ? G.boxEnd[0] - 1 : G.boxStart[0]);
            const float y = float((c & 2) ? G.boxEnd[1] - 1 : G.boxStart[1]);
            const float z = float((c & 4) ? G.boxEnd[2] - 1 : G.boxStart[2]);
            tmax = fmaxf(tmax, sampled_plane(g, inside, x, y, z));
        }
        const int tailEnd = min(G.passive, int(floorf(tmax)) + 2);
        G.tailCount = max(0, tailEnd - G.tailStart);
        G.pad = 0;
        s_gs[j] = G;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int base = 0;
        for (int j = 0; j < m; ++j) {
            s_gs[j].tailBase = base;
            base += s_gs[j].tailCount;
        }
        s_total = base;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < m; j += blockDim.x) {
        const GroupScalars G = s_gs[j];
        for (int q = 0; q < G.tailCount && G.tailBase + q < tailCap; ++q) {
            tailPlaneGroup[G.tailBase + q] = j;
            tailPlaneK[G.tailBase + q] = G.tailStart + q;
        }
        gs[j] = G;
    }
    for (int l = threadIdx.x; l < L; l += blockDim.x) {
        const int j = s_group[l];
        ls[l].group = j;
        ls[l].mainEnd = j >= 0 ? s_gs[j].tailStart : inside;
    }
    // union box and tail totals: one wave reduces over the groups
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        int mp = INT_MIN, me = inside;
        int st0 = INT_MAX, st1 = INT_MAX, st2 = INT_MAX, en0 = INT_MIN, en1 = INT_MIN, en2 = INT_MIN;
        int mi0 = -1, mi1 = -1, mi2 = -1;
        for (int j = lane; j < m; j += 64) {
            const GroupScalars &G = s_gs[j];
            mp = max(mp, G.passive);
            me = max(me, G.tailStart);
            st0 = min(st0, G.boxStart[0]); st1 = min(st1, G.boxStart[1]); st2 = min(st2, G.boxStart[2]);
            en0 = max(en0, G.boxEnd[0]); en1 = max(en1, G.boxEnd[1]); en2 = max(en2, G.boxEnd[2]);
            mi0 = max(mi0, G.boxMax[0]); mi1 = max(mi1, G.boxMax[1]); mi2 = max(mi2, G.boxMax[2]);
        }
        mp = wave_max_uniform_i(mp); me = wave_max_uniform_i(me);  // wave 0, every lane active
        st0 = wave_min_uniform_i(st0); st1 = wave_min_uniform_i(st1); st2 = wave_min_uniform_i(st2);
        en0 = wave_max_uniform_i(en0); en1 = wave_max_uniform_i(en1); en2 = wave_max_uniform_i(en2);
        mi0 = wave_max_uniform_i(mi0); mi1 = wave_max_uniform_i(mi1); mi2 = wave_max_uniform_i(mi2);
        if (lane == 0) {
            const int total = s_total;
            bs->numGroups = m;
            bs->tailPlanes = min(total, tailCap);
            bs->overflow = total > tailCap ? 1 : 0;
            bs->maxPassive = mp;
            bs->mainEndMax = me;
            bs->activeLayers = m > 0 ? 1 : 0;
            bs->boxStart[0] = m > 0 ? st0 : 0; bs->boxStart[1] = m > 0 ? st1 : 0; bs->boxStart[2] = m > 0 ? st2 : 0;
            bs->boxEnd[0] = m > 0 ? en0 : 0; bs->boxEnd[1] = m > 0 ? en1 : 0; bs->boxEnd[2] = m > 0 ? en2 : 0;
            bs->boxMax[0] = mi0; bs->boxMax[1] = mi1; bs->boxMax[2] = mi2;
        }
    }
}
struct LayerFinalizeK {
    static constexpr int NT = 256;
    struct Args {
        FanGeom g;
        LayerScalars *ls;
        GroupScalars *gs;
        BeamScalars *bs;
        int *tailPlaneGroup;
        int *tailPlaneK;
        int tailCap;
        int *planeRM;
    };
    static __device__ __forceinline__ void run(const Args &a, const uint3 b, const uint3 d) { layer_finalize_body(b, d, a.g, a.ls, a.gs, a.bs, a.tailPlaneGroup, a.tailPlaneK, a.tailCap, a.planeRM); }
};


// The halo grid reuses the beam's passive steps; its groups, boxes and tails are its own.
// Copies the primary scalars and clears the grid's radius histograms and class counters.
__device__ __forceinline__ void seed_grid_body(const uint3 blockIdx, const uint3 gridDim, int numLayers, const LayerScalars *__restrict__ ls, const BeamScalars *__restrict__ bs,
                            LayerScalars *__restrict__ lsOut, BeamScalars *__restrict__ bsOut, int *__restrict__ hist,
                            int *__restrict__ classCount) {
    for (int l = threadIdx.x; l < numLayers; l += blockDim.x) lsOut[l] = ls[l];
    for (int i = threadIdx.x; i < numLayers * (kMaxRadDev + 2); i += blockDim.x) hist[i] = 0;
    if (threadIdx.x < 16) classCount[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        BeamScalars b = *bs;
        b.radOverflow = 0;
        *bsOut = b;
    }
}
struct SeedGridK {
    static constexpr int NT = 256;
    struct Args {
        int numLayers;
        const LayerScalars *ls;
        const BeamScalars *bs;
        LayerScalars *lsOut;
        BeamScalars *bsOut;
        int *hist;
        int *classCount;
    };
    static __device__ __forceinline__ void run(const Args &a, const uint3 b, const uint3 d) { seed_grid_body(b, d, a.numLayers, a.ls, a.bs, a.lsOut, a.bsOut, a.hist, a.classCount); }
};


// (0, +inf) over a halo (IDD, rSigmaEff) field: cells no spot ray writes stay empty.
__device__ __forceinline__ void clear_pairs_body(const uint3 blockIdx, const uint3 gridDim, float2 *__restrict__ v, size_t n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        v[i] = make_float2(0.0f, PBS_INF);
}
struct ClearPairsK {
    static constexpr int NT = 256;
    struct Args {
        float2 *v;
        size_t n;
    };
    static __device__ __forceinline__ void run(const Args &a, const uint3 b, const uint3 d) { clear_pairs_body(b, d, a.v, a.n); }
};


// ---------------------------------------------------------------------------------------
// tile radii, bins and work items
// ---------------------------------------------------------------------------------------
// one wave per (step, tile); loops over layers. code = rad | live << 8, -1 = not computed
__device__ __forceinline__ void tiles_body(const uint3 blockIdx, const uint3 gridDim, FanGeom g, const float2 *__restrict__ iddSigma,
                                               const LayerScalars *__restrict__ ls, const BeamScalars *__restrict__ bs,
                                               int *__restrict__ tileRad, int *__restrict__ hist) {
    const int ntx = g.sx / 32, nty = g.sy / 8, nt = ntx * nty;
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int k = blockIdx.y;
    const int lane = threadIdx.x & 63;
    if (t >= nt) return;
    const int tx = t % ntx, ty = t / ntx;
    const int inside = bs->beamFirstInside;
    for (int l = 0; l < g.numLayers; ++l) {
        const size_t code_idx = (size_t(l) * g.K + k) * nt + t;
        const int pas = ls[l].passive;
        if (k < inside || k >= pas) {
            if (lane == 0) tileRad[code_idx] = -1;
            continue;
        }
        float m = __builtin_nanf("");
        bool any = false;
        const float2 *base = iddSigma + (size_t(l) * g.K + k) * g.sx * g.sy;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int p = lane + 64 * q;
            const int x = tx * 32 + (p & 31), y = ty * 8 + (p >> 5);
            float2 v = make_float2(0.0f, PBS_INF);
            if (x < g.rx && y < g.ry) v = base[size_t(y) * g.sx + x];
            m = fminf(m, v.y);
            any |= v.x > 0.0f;
        }
        m = wave_min(m);
        const bool live = __any(any);
        if (lane == 0) {
            const int r = tile_radius(m);
            atomicAdd(&hist[l * (kMaxRadDev + 2) + r], 1);
            tileRad[code_idx] = r | (live ? 256 : 0);
        }
    }
}
struct TilesK {
    static constexpr int NT = 256;
    struct Args {
        FanGeom g;
        const float2 *iddSigma;
        const LayerScalars *ls;
        const BeamScalars *bs;
        int *tileRad;
        int *hist;
    };
    static __device__ __forceinline__ void run(const Args &a, const uint3 b, const uint3 d) { tiles_body(b, d, a.g, a.iddSigma, a.ls, a.bs, a.tileRad, a.hist); }
};


// overflowFlag (halo grid, complete_superposition.cu:558-620): a tile past the largest radius
// sets the sticky flag the host reports as an error; the primary grid (nullptr) only clamps
// it (413-463). Both clamp, so every launch stays within the largest radius.
// One layer's launch radii from its radius histogram `hl` (kMaxRadDev + 2 counts) into `effL`.
__device__ __forceinline__ void layer_bins(const int *hl, int *__restrict__ effL, int *__restrict__ overflowFlag) {
    // every loop has a fixed trip count and is unrolled, so h and eff stay in registers (a loop
    // bound by layerMax would index them at run time and put them in scratch memory)
    int h[kMaxRadDev + 2];
#pragma unroll
    for (int i = 0; i < kMaxRadDev + 2; ++i) h[i] = hl[i];
    if (overflowFlag && h[kMaxRadDev + 1] > 0) *overflowFlag = 1;
    h[kMaxRadDev] += h[kMaxRadDev + 1];  // primary overflow clamps into the largest radius
    h[kMaxRadDev + 1] = 0;
    int layerMax = 0;
#pragma unroll
    for (int i = 0; i < kMaxRadDev + 2; ++i) layerMax = h[i] > 0 ? i : layerMax;
    int eff[kMaxRadDev + 2];
#pragma unroll
    for (int i = 0; i < kMaxRadDev + 2; ++i) eff[i] = i;
    eff[kMaxRadDev + 1] = kMaxRadDev;
    // from the largest radius down: bins are batched upwards until kMinTilesDev tiles
    int batched = 0, rec = layerMax;
#pragma unroll
    for (int r = kMaxRadDev; r > 0; --r) {
        if (r <= layerMax) {
            batched += h[r];
            eff[r] = rec;
            if (batched >= kMinTilesDev) { rec = r - 1; batched = 0; }
        }
    }
#pragma unroll
    for (int i = 0; i < kMaxRadDev + 2; ++i) effL[i] = eff[i];
}
__device__ __forceinline__ void bins_body(const uint3 blockIdx, const uint3 gridDim, int numLayers, const int *__restrict__ hist, int *__restrict__ effRad,
                       int *__restrict__ overflowFlag) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= numLayers) return;
    layer_bins(hist + l * (kMaxRadDev + 2), effRad + l * (kMaxRadDev + 2), overflowFlag);
}
struct BinsK {
    static constexpr int NT = 64;
    struct Args {
        int numLayers;
        const int *hist;
        int *effRad;
        int *overflowFlag;
    };
    static __device__ __forceinline__ void run(const Args &a, const uint3 b, const uint3 d) { bins_body(b, d, a.numLayers, a.hist, a.effRad, a.overflowFlag); }
};

// The layer finalize, the per-layer radius histograms and the launch radii in one launch: work
// group l < L histograms layer l's tile codes over [beamFirstInside, passive) (its passive step
// computed as the finalize computes it) and bins them (thread 0); work-group L runs the finalize
// (groups, boxes, tail planes). None of them reads what another writes.
struct FinalizeTilesK {
    static constexpr int NT = 256;
    struct Args {
        LayerFinalizeK::Args fin;
        const int *tileCode;
        int *hist;
        int *effRad;
        int *overflowFlag;
    };
    static __device__ __forceinline__ void run(const Args &a, const uint3 b, const uint3 d) {
        const int L = a.fin.g.numLayers;
        if (int(b.x) == L) {
            LayerFinalizeK::run(a.fin, uint3{0u, 0u, 0u}, uint3{1u, 1u, 1u});
            return;
        }
        __shared__ int s_h[kMaxRadDev + 2];
        tile_hist_into(b, a.fin.g, a.tileCode, a.fin.ls, a.fin.bs, s_h);
        const int l = b.x;
        if (threadIdx.x < kMaxRadDev + 2) a.hist[l * (kMaxRadDev + 2) + threadIdx.x] = s_h[threadIdx.x];
        if (threadIdx.x == 0) layer_bins(s_h, a.effRad + l * (kMaxRadDev + 2), a.overflowFlag);
        (void)d;
    }
};


__device__ __forceinline__ int radius_class(int R) {
    // launch classes: accumulator layouts of radius 7, 11, 15 (2RM + 1 lanes fill a group of
    // 16, 32, 32), then 24, 32
    return R <= 7 ? 0 : R <= 11 ? 1 : R <= 15 ? 2 : R <= 24 ? 3 : 4;
}
__device__ __forceinline__ int class_radius(int c) {
    constexpr int rm[kNumClasses] = {7, 11, 15, 24, 32};
    return rm[c];
}

__device__ __forceinline__ bool in_tail(const GroupScalars &G, int k) {
    return k >= G.tailStart && k < G.tailStart + G.tailCount;
}

// Appends a work unit to its class list; lanes of a wave that append to the same class share
// one counter update.
__device__ __forceinline__ void append_unit(int *__restrict__ classCount, int2 *__restrict__ classList, int maxUnits,
                                            int c, int2 unit) {
    const uint64_t active = __ballot(1);
    for (int cc = 0; cc < kNumClasses; ++cc) {
        const uint64_t peers = __ballot(c == cc);
        if (!peers) continue;
        const int leader = __builtin_ctzll(peers);
        const int lane = __lane_id();
        int base = 0;
        if (lane == leader) base = atomicAdd(&classCount[cc], __popcll(peers));
        base = __builtin_amdgcn_readlane(base, leader);  // leader is uniform
        if (c == cc) {
            const int pos = base + __popcll(peers & ((uint64_t(1) << lane) - 1));
            if (pos < maxUnits) classList[size_t(cc) * maxUnits + pos] = unit;
        }
    }
    (void)active;
}

// One thread per (step, tile): the item's launch class from the largest launch radius of its
// live layers, the block layout radius of its main block and group tail blocks (-1 when a
// block is empty), and one work unit (slot, block) per non-empty block in its class list.
// The layer scalars, launch radii and group tail ranges are staged in LDS and the tile codes
// are loaded eight layers at a time, so the loop's global loads do not wait on each other.
__device__ __forceinline__ void items_body(const uint3 blockIdx, const uint3 gridDim, FanGeom g, const int *__restrict__ tileRad,
                                               const int *__restrict__ effRad, const LayerScalars *__restrict__ ls,
                                               const GroupScalars *__restrict__ gs, const BeamScalars *__restrict__ bs,
                                               int *__restrict__ itemR, int *__restrict__ tailItemR,
                                               int *__restrict__ classCount, int2 *__restrict__ classList,
                                               int maxUnits, int *__restrict__ planeRM,
                                               int *__restrict__ slotLayers) {
    extern __shared__ int s_items[];
    const int L = g.numLayers, m = bs->numGroups;
    int *s_pass = s_items, *s_group = s_pass + L, *s_mainEnd = s_group + L;
    int *s_eff = s_mainEnd + L;                 // [L][kMaxRadDev + 2]
    int *s_tail = s_eff + L * (kMaxRadDev + 2);  // [m][3]: tailStart, tailCount, tailBase
    for (int i = threadIdx.x; i < L; i += blockDim.x) {
        s_pass[i] = ls[i].passive;
        s_group[i] = ls[i].group;
        s_mainEnd[i] = ls[i].mainEnd;
    }
    for (int i = threadIdx.x; i < L * (kMaxRadDev + 2); i += blockDim.x) s_eff[i] = effRad[i];
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        s_tail[3 * i] = gs[i].tailStart;
        s_tail[3 * i + 1] = gs[i].tailCount;
        s_tail[3 * i + 2] = gs[i].tailBase;
    }
    __syncthreads();
    const int nt = (g.sx / 32) * (g.sy / 8);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    // waves past the last tile leave; lanes past it stay (inactive) for the wave reductions
    if (t - int(threadIdx.x & 63) >= nt) return;
    const bool tileIn = t < nt;
    auto inTail = [&](int j) { return k >= s_tail[3 * j] && k < s_tail[3 * j] + s_tail[3 * j + 1]; };
    int Rmax = -1;
    bool mainLive = false;
    uint64_t tailLive[kMaxLayersDev / 64] = {0, 0, 0, 0};  // groups with a live tail layer here
    const int inside = bs->beamFirstInside;
    const int slot = k * nt + t;
    // the slot's live layers in layer order, each l | R << 8 | (block + 1) << 16 (block -1 =
    // main BEV, j = layer group j's tail), terminated by -1 when fewer than numLayers
    int *list = slotLayers + size_t(slot) * L;
    int n = 0;
    for (int l0 = 0; l0 < L; l0 += 8) {
        int code[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            code[u] = (tileIn && l0 + u < L) ? tileRad[(size_t(l0 + u) * g.K + k) * nt + t] : -1;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int l = l0 + u;
            if (l >= L || k < inside || k >= s_pass[l]) continue;  // steps the layer does not superpose
            if (code[u] < 0 || !(code[u] & 256)) continue;
            const int j = s_group[l];
            const bool isMain = k < s_mainEnd[l];
            if (!isMain && !(j >= 0 && inTail(j))) continue;
            mainLive |= isMain;
            if (!isMain) tailLive[j >> 6] |= uint64_t(1) << (j & 63);
            const int R = s_eff[l * (kMaxRadDev + 2) + (code[u] & 255)];
            Rmax = max(Rmax, R);
            list[n++] = l | (R << 8) | ((isMain ? 0 : j + 1) << 16);
        }
    }
    if (tileIn && n < L) list[n] = -1;
    const int c = Rmax >= 0 ? radius_class(Rmax) : 0;
    const int RM = Rmax >= 0 ? class_radius(c) : -1;
    // the plane's largest block layout radius: one atomic per wave
    const int lane = threadIdx.x & 63;
    if (tileIn) itemR[slot] = mainLive ? RM : -1;
    if (tileIn && mainLive) append_unit(classCount, classList, maxUnits, c, make_int2(slot, -1));
    {
        const int pm = wave_max_uniform_i(mainLive ? RM : -1);  // every lane is active here
        if (lane == 0 && pm >= 0) atomicMax(&planeRM[k], pm);
    }
    for (int j = 0; j < m; ++j) {
        if (!inTail(j)) continue;  // uniform: k is the block's step
        const int q = s_tail[3 * j + 2] + (k - s_tail[3 * j]);
        if (q >= bs->tailPlanes) continue;
        const bool live = tileIn && ((tailLive[j >> 6] >> (j & 63)) & 1);
        if (tileIn) tailItemR[size_t(q) * nt + t] = live ? RM : -1;
        if (live) append_unit(classCount, classList, maxUnits, c, make_int2(slot, j));
        const int pm = wave_max_uniform_i(live ? RM : -1);
        if (lane == 0 && pm >= 0) atomicMax(&planeRM[g.K + q], pm);
    }
}
struct ItemsK {
    static constexpr int NT = 256;
    struct Args {
        FanGeom g;
        const int *tileRad;
        const int *effRad;
        const LayerScalars *ls;
        const GroupScalars *gs;
        const BeamScalars *bs;
        int *itemR;
        int *tailItemR;
        int *classCount;
        int2 *classList;
        int maxUnits;
        int *planeRM;
        int *slotLayers;
    };
    static __device__ __forceinline__ void run(const Args &a, const uint3 b, const uint3 d) { items_body(b, d, a.g, a.tileRad, a.effRad, a.ls, a.gs, a.bs, a.itemR, a.tailItemR, a.classCount, a.classList, a.maxUnits, a.planeRM, a.slotLayers); }
};


// ---------------------------------------------------------------------------------------
// superposition
// ---------------------------------------------------------------------------------------
// Superposition of a 2-row slice of a tile by one wave.
//
// Lane c*64 + lane owns output column X = c*64 + lane - RM (tile coordinates; the block spans
// X in [-RM, 32 + RM)) and keeps the 2 + 2RM output rows its slice reaches in registers.
// The slice's 64 source pixels are staged in the wave's LDS table as [d, e_0 .. e_R]. A sweep
// visits each source row in 2R+1 phases; in phase phi a lane takes the one source
// x_s = phi (mod 2R+1) of the row with |X - x_s| <= R, so the sources of one phase have
// disjoint footprints and all of them are applied by one pass of 2R+1 FMAs per lane
// (d e[|X - x_s|] e[|dy|] into row y_s + dy). Every output is a sum in a fixed order: the
// result is bitwise repeatable.
constexpr int pad4(int n) { return (n + 3) / 4 * 4; }

template <int RM>
struct SliceAcc {
    static constexpr int NC = (32 + 2 * RM + 63) / 64;
    static constexpr int NR = 2 + 2 * RM;
    float v[NC][NR];
};

template <int R, int RM>
__device__ __forceinline__ void sweep_slice(SliceAcc<RM> &acc, const float *__restrict__ tab, unsigned rowLive,
                                            int lane) {
    constexpr int NC = SliceAcc<RM>::NC;
    constexpr int EWp = pad4(RM + 2);
    constexpr int S = 2 * R + 1;
    constexpr int NE4 = (R + 2 + 3) / 4;  // float4 reads covering [d, e_0 .. e_R]
    int rem[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const int u = c * 64 + lane - RM + R + S * 64;  // X + R, made positive
        rem[c] = u % S;
    }
#pragma unroll 1
    for (int phi = 0; phi < S; ++phi) {
        int base[NC], adx[NC];
        bool valid[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int xs = c * 64 + lane - RM + R - rem[c];
            valid[c] = xs >= 0 && xs < 32;
            base[c] = (valid[c] ? xs : 0) * EWp;
            adx[c] = rem[c] > R ? rem[c] - R : R - rem[c];
        }
#pragma unroll
        for (int yy = 0; yy < 2; ++yy) {
            if (!(rowLive & (1u << yy))) continue;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const float *src = tab + yy * 32 * EWp + base[c];
                float E[NE4 * 4];
#pragma unroll
                for (int q = 0; q < NE4; ++q) {
                    const float4 v = *reinterpret_cast<const float4 *>(src + 4 * q);
                    E[4 * q] = v.x; E[4 * q + 1] = v.y; E[4 * q + 2] = v.z; E[4 * q + 3] = v.w;
                }
                const float a = valid[c] ? E[0] * src[1 + adx[c]] : 0.0f;
#pragma unroll
                for (int i = -R; i <= R; ++i) acc.v[c][yy + RM + i] += a * E[1 + (i < 0 ? -i : i)];
            }
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) rem[c] = rem[c] == 0 ? S - 1 : rem[c] - 1;
    }
}

template <int R, int RM>
__device__ __forceinline__ void sweep_dispatch(int r, SliceAcc<RM> &acc, const float *tab, unsigned rowLive,
                                               int lane) {
    if constexpr (R <= RM) {
        if (r == R)
            sweep_slice<R, RM>(acc, tab, rowLive, lane);
        else
            sweep_dispatch<R + 1, RM>(r, acc, tab, rowLive, lane);
    }
}

// One work-group of 4 waves per work unit (step, tile, block) of radius class RM; the block
// is the main BEV block or one layer group's tail block. Wave w takes source rows 2w, 2w+1
// of every live layer of the block (complete_superposition.cu:280-330: per-pixel erf
// differences e0 = erf(r/2), e_i = (erf(r(i+1/2)) - erf(r(i-1/2)))/2, outer product
// d e_x e_y), so all waves work on every layer; the four partial blocks are added in wave
// order in LDS and written to the unit's slot as a (8 + 2RM) x (32 + 2RM) block. Fixed
// orders throughout: bitwise repeatable.
template <int RM>
__device__ __forceinline__ void superpose_body(const uint3 blockIdx, const uint3 gridDim, FanGeom g, const float2 *__restrict__ iddSigma,
                                                   const int *__restrict__ tileRad, const int *__restrict__ effRad,
                                                   const LayerScalars *__restrict__ ls,
                                                   const GroupScalars *__restrict__ gs,
                                                   const BeamScalars *__restrict__ bs,
                                                   const int *__restrict__ classCount,
                                                   const int2 *__restrict__ classList, int cls, int maxUnits,
                                                   float *__restrict__ mainSlots, float *__restrict__ tailSlots) {
    constexpr int NC = SliceAcc<RM>::NC, NR = SliceAcc<RM>::NR;
    constexpr int EWp = pad4(RM + 2);
    constexpr int BW = 32 + 2 * RM, BH = 8 + 2 * RM;
    constexpr int LJ = kMaxLayersDev / 64;
    __shared__ __align__(16) float s_tab[4][64 * EWp];  // per wave: its 2 source rows x 32 pixels
    __shared__ float s_B[BH * BW];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    float *tab = s_tab[w];
    const int nt = (g.sx / 32) * (g.sy / 8), ntx = g.sx / 32;
    const int count = min(classCount[cls], maxUnits);
    const int2 *list = classList + size_t(cls) * maxUnits;
    for (int u = blockIdx.x; u < count; u += gridDim.x) {
        const int2 unit = list[u];
        const int slot = unit.x, b = unit.y;
        const int k = slot / nt, t = slot % nt;
        const int tx = t % ntx, ty = t / ntx;
        // this block's layers, one layer per lane, loaded together
        uint64_t lmask[LJ];
        int radv[LJ];
#pragma unroll
        for (int j = 0; j < LJ; ++j) {
            const int l = 64 * j + lane;
            bool in = false;
            int rad = 0;
            if (l < g.numLayers && k >= bs->beamFirstInside && k < ls[l].passive) {
                const int code = tileRad[(size_t(l) * g.K + k) * nt + t];
                if (code >= 0 && (code & 256)) {
                    const LayerScalars L = ls[l];
                    in = b < 0 ? (k < L.mainEnd) : (k >= L.mainEnd && L.group == b);
                    if (in) rad = effRad[l * (kMaxRadDev + 2) + (code & 255)];
                }
            }
            lmask[j] = __ballot(in);
            radv[j] = rad;
        }
        SliceAcc<RM> acc;
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int q = 0; q < NR; ++q) acc.v[c][q] = 0.0f;
        // this lane's source pixel: column lane & 31 of row 2w + (lane >> 5)
        const int gx = tx * 32 + (lane & 31), gy = ty * 8 + 2 * w + (lane >> 5);
        const bool pixIn = gx < g.rx && gy < g.ry;
        const size_t pixOff = size_t(gy) * g.sx + gx;
        float *e = tab + ((lane >> 5) * 32 + (lane & 31)) * EWp;
#pragma unroll
        for (int j = 0; j < LJ; ++j) {
            uint64_t msk = lmask[j];
            while (msk) {
                const int ln = __builtin_ctzll(msk);
                msk &= msk - 1;
                const int l = 64 * j + ln;
                const int R = __builtin_amdgcn_readlane(radv[j], ln);
                float d = 0.0f, r = PBS_INF;
                if (pixIn) {
                    const float2 v = iddSigma[(size_t(l) * g.K + k) * g.sy * g.sx + pixOff];
                    d = v.x;
                    r = v.y;
                }
                const uint64_t live = __ballot(d != 0.0f);
                const unsigned rowLive = (unsigned(uint32_t(live) != 0u)) | (unsigned(uint32_t(live >> 32) != 0u) << 1);
                if (!rowLive) continue;  // uniform
                e[0] = d;
                {
                    float en = erff(r * 0.5f), eo = -en;
                    for (int i = 0; i <= R; ++i) {
                        e[1 + i] = 0.5f * (en - eo);
                        eo = en;
                        if (i < R) en = erff(r * (float(i) + 1.5f));
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                sweep_dispatch<0, RM>(R, acc, tab, rowLive, lane);
                // the next staging overwrites the table: keep it after these reads
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        }
        // partial blocks added in wave order (wave w's rows start at block row 2w)
        for (int i = threadIdx.x; i < BH * BW; i += 256) s_B[i] = 0.0f;
        __syncthreads();
        for (int ww = 0; ww < 4; ++ww) {
            if (w == ww) {
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    const int col = c * 64 + lane;
                    if (col < BW) {
#pragma unroll
                        for (int q = 0; q < NR; ++q) s_B[(2 * w + q) * BW + col] += acc.v[c][q];
                    }
                }
            }
            __syncthreads();
        }
        float *out;
        if (b < 0) {
            out = mainSlots + size_t(slot) * kSlotFloats;
        } else {
            const int q = gs[b].tailBase + (k - gs[b].tailStart);
            out = (q < bs->tailPlanes) ? tailSlots + (size_t(q) * nt + t) * kSlotFloats : nullptr;
        }
        if (out)
            for (int i = threadIdx.x; i < BH * BW; i += 256) out[(i / BW) * kSlotPitch + (i % BW)] = s_B[i];
        __syncthreads();
    }
}
template <int RM>
struct SuperposeK {
    static constexpr int NT = 256;
    static constexpr bool kUnitQueue = true;  // batched as a queue of the beams' units
    struct Args {
        FanGeom g;
        const float2 *iddSigma;
        const int *tileRad;
        const int *effRad;
        const LayerScalars *ls;
        const GroupScalars *gs;
        const BeamScalars *bs;
        const int *classCount;
        const int2 *classList;
        int cls;
        int maxUnits;
        float *mainSlots;
        float *tailSlots;
    };
    static __device__ __forceinline__ void run(const Args &a, const uint3 b, const uint3 d) { superpose_body<RM>(b, d, a.g, a.iddSigma, a.tileRad, a.effRad, a.ls, a.gs, a.bs, a.classCount, a.classList, a.cls, a.maxUnits, a.mainSlots, a.tailSlots); }
};


// ---------------------------------------------------------------------------------------
// Superposition with lanes on output rows (radius classes up to 16)
// ---------------------------------------------------------------------------------------
// Table row of one source: row[1 + m] = e_m, the pixel weights of a Gaussian of reciprocal
// width r (complete_superposition.cu:280-300: e_0 = erf(r/2), e_m = (erf(r(m+1/2)) - erf(r(m-1/2)))/2),
// for m <= R and 0 past R. erf(x) = 1 - T(x) with the tail T(x) = t P(t) exp(-x^2), t = 1/(1 + p x)
// (Abramowitz & Stegun 7.1.26, |error| <= 1.5e-7), so e_0 = 1 - T_0 and e_m = (T_{m-1} - T_m)/2: a
// difference of tails, which keeps the small weights' relative accuracy. exp(-x_m^2) at x_m =
// r(m + 1/2) follows x_{m+1}^2 - x_m^2 = 2 r^2 (m + 1): g_{m+1} = g_m h_m with h_m = c^(m+1),
// c = exp(-2 r^2), so a row takes two exponentials and R + 1 reciprocals. r = inf (a dead pixel)
// gives e_0 = 1 and zeros; r = 0 gives zeros.
template <int EW>
__device__ __forceinline__ void gauss_row(float (&row)[EW], float r, int R) {
    const float r2 = r * r;
    const float c = __builtin_amdgcn_exp2f(-2.88539008f * r2);  // exp(-2 r^2)
    float g = __builtin_amdgcn_exp2f(-0.360673760f * r2);       // exp(-r^2 / 4)
    float h = c, tPrev = 0.0f;
#pragma unroll
    for (int m = 0; m < EW - 1; ++m) {
        float e = 0.0f;
        if (m <= R) {
            const float t = rcp_hw(1.0f + 0.3275911f * (r * (float(m) + 0.5f)));
            const float T =
                g * (t * (0.254829592f + t * (-0.284496736f + t * (1.421413741f + t * (-1.453152027f + t * 1.061405429f)))));
            e = m == 0 ? 1.0f - T : 0.5f * (tPrev - T);
            tPrev = T;
            g *= h;
            h *= c;
        }
        row[1 + m] = e;
    }
}

// Work-group barrier for LDS hand-offs only: waits for this wave's LDS accesses, not for its
// global loads, so loads issued ahead (the next layer's pixels) stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0); vmcnt and expcnt left at their maximum
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// Launch shape of the row-lane superposition of class RM. The vertical weight of a source row
// is e_|dy|, the same for the output rows y_s - q and y_s + q: a group of G >= RM + 1 lanes holds
// one source row y_s and one half of its sources (x_s in [16h, 16h + 16)), lane q sweeping the
// horizontal rows of |dy| = q once for both output rows. Sixteen groups (eight source rows times
// two halves; a wave holds groups of one half) make a work-group of 16G threads. Each source's
// table row is [d, e_0 .. e_R, 0 ...] (EW floats).
template <int RM>
struct RowsShape {
    static constexpr int BH = 8 + 2 * RM, BW = 32 + 2 * RM;
    static constexpr int HW = 16 + 2 * RM;  // the columns one half's sources reach
    static constexpr int G = RM + 1 <= 8 ? 8 : (RM + 1 <= 16 ? 16 : (RM + 1 <= 32 ? 32 : 64));
    static constexpr int NT = 16 * G;
    static constexpr int EW = pad4(RM + 3);
    // The table rows of a source row's 32 columns go out as 16-byte stores, which the LDS serves 8
    // lanes (128 bytes) at a time: 8 consecutive columns must start on distinct 4-bank groups
    // modulo 32 banks. A column stride that is an odd multiple of 4 floats does that (12, 20: the
    // starts of any 8 aligned columns are distinct modulo 32 banks, of any 16 modulo 64). A stride
    // of 16 would start them on 2 groups only (4-way conflicts): an EW of 16 gets a stride of 20.
    static constexpr int CS = EW % 8 == 0 ? EW + 4 : EW;
    static constexpr int SPAN = 32 * CS;
    // source-row pitch: SPAN rounded up to G modulo 64 floats puts the 64 / G groups of a wave at
    // bank offsets 0, G, 2G, ...: each group's per-lane reads E[1 + v] (v <= RM < G) then fall
    // on banks no other group of the wave touches
    static constexpr int RP = SPAN + ((G - SPAN % 64) % 64 + 64) % 64;
    static __device__ __forceinline__ int col_off(int col) { return col * CS; }
    static constexpr int EBUF = 8 * RP;  // one layer's table
    // after the last layer: four groups' rows at a time. The pitch is odd, so the lanes of a wave
    // write distinct banks. The block sum reads 32 lanes at a time along block rows, and a
    // half-wave often spans two rows |dy| apart by one: at RM = 11 the pitch 39 puts such rows'
    // columns on common banks (77 conflicted cycles per unit), 43 none (nor 31 and 47 at RM 7, 15)
    static constexpr int HWp = RM == 11 ? 43 : (HW | 1);
    static constexpr int ACC = 4 * G * HWp;
    static constexpr int OUT = BH * BW;
    static constexpr int LDS = EBUF > ACC ? EBUF : ACC;
    static constexpr int SPT = (256 + NT - 1) / NT;      // sources staged per thread
    static constexpr int OPT = (OUT + NT - 1) / NT;      // output pixels per thread
    static constexpr int RLO = RM - 7 > 0 ? RM - 7 : 0;  // smallest compiled sweep radius
    static_assert(RM + 1 <= G, "a group holds the vertical offsets 0 .. RM");
    static_assert(8 * G % 64 == 0, "a wave holds groups of one half only");
};

// One source row of one layer into this lane's row: for every source x_s, a = d e_v with v = |dy|
// clamped to R+1 (a zero entry), then 2R+1 FMAs along the row, acc[x_s + dx + RM] += a e_|dx|. The sources go in chunks of four; a chunk where no source row of the wave has a live
// pixel (bit x_s of wm, the OR of the wave's live masks, uniform over the wave) is skipped by a
// scalar branch: its sources would add exact zeros. Within a chunk the sweep is branch-free. RP_
// is the radius compiled; entries past the layer's own radius are zero, so a smaller layer
// radius adds exact zeros.
template <int RP_, int RM, int EW, int H>
__device__ __forceinline__ void rows_sweep(float (&acc)[16 + 2 * RM], const float *__restrict__ rowE, int v,
                                           uint32_t wm) {
    constexpr int NE4 = (RP_ + 2 + 3) / 4;
    using S = RowsShape<RM>;
#pragma unroll
    for (int x0 = 16 * H; x0 < 16 * H + 16; x0 += 4) {
        if (!((wm >> x0) & 0xfu)) continue;  // uniform: wm is a scalar
#pragma unroll
        for (int xs = x0; xs < x0 + 4; ++xs) {
            const float *E = rowE + S::col_off(xs);
            float e[4 * NE4];
#pragma unroll
            for (int q = 0; q < NE4; ++q) {
                const float4 t4 = *reinterpret_cast<const float4 *>(E + 4 * q);
                e[4 * q] = t4.x; e[4 * q + 1] = t4.y; e[4 * q + 2] = t4.z; e[4 * q + 3] = t4.w;
            }
            const float a = e[0] * E[1 + v];
#pragma unroll
            for (int dx = -RP_; dx <= RP_; ++dx) {
                const int c = xs - 16 * H + dx + RM;
                acc[c] = fmaf(a, e[1 + (dx < 0 ? -dx : dx)], acc[c]);
            }
        }
    }
}

template <int RP_, int RM, int EW, int H>
__device__ __forceinline__ void rows_dispatch(int rp, float (&acc)[16 + 2 * RM], const float *rowE, int v,
                                              uint32_t wm) {
    if constexpr (RP_ <= RM) {
        if (rp == RP_)
            rows_sweep<RP_, RM, EW, H>(acc, rowE, v, wm);
        else
            rows_dispatch<RP_ + 1, RM, EW, H>(rp, acc, rowE, v, wm);
    }
}

// Work-group per work unit (step, tile, block) of class RM (complete_superposition.cu:280-330:
// e_0 = erf(r/2), e_i = (erf(r(i+1/2)) - erf(r(i-1/2)))/2, outer product d e_x e_y). Lane q of
// group (h, g) keeps, across all layers of the unit, the 16 + 2RM columns that the sources
// [16h, 16h + 16) of source row g reach, weighted by e_q: their contribution to block rows
// g + RM - q and g + RM + q alike. Per layer the work-group stages the tile's 256 table rows once
// (one barrier before the sweep, one after it; the next layer's pixels are loaded during the
// sweep); every source is then applied by all lanes of its group at once, so the horizontal
// weights are shared reads and the register index is static. At the end each block pixel adds
// the sixteen groups' rows in group order: bitwise repeatable.
template <int RM>
__device__ __forceinline__ void superpose_rows_body(const uint3 blockIdx, const uint3 gridDim, FanGeom g,
                                                    const float2 *__restrict__ iddSigma,
                                                    const int *__restrict__ tileRad, const int *__restrict__ effRad,
                                                    const LayerScalars *__restrict__ ls,
                                                    const GroupScalars *__restrict__ gs,
                                                    const BeamScalars *__restrict__ bs,
                                                    const int *__restrict__ classCount,
                                                    const int2 *__restrict__ classList, int cls, int maxUnits,
                                                    const int *__restrict__ slotLayers, float *__restrict__ mainSlots,
                                                    float *__restrict__ tailSlots) {
    using S = RowsShape<RM>;
    constexpr int BW = S::BW, HW = S::HW, G = S::G, NT = S::NT, EW = S::EW, RP = S::RP, HWp = S::HWp;
    constexpr int SPT = S::SPT, OPT = S::OPT;
    __shared__ __align__(16) float s_lds[S::LDS];
    __shared__ uint32_t s_live[8];
    __shared__ int s_list[kMaxLayersDev];  // the unit's layers: l | R << 16
    __shared__ int s_n;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int gi = tid / G, q = tid % G;  // group, vertical offset |dy|
    const int grp = gi % 8;  // source row
    // half of the sources: the same for the whole wave, and read as a scalar so that the chunk
    // tests on the live mask below are scalar branches (a per-lane condition would make them
    // divergent branches, with the accumulators copied around every chunk)
    const int half = __builtin_amdgcn_readfirstlane(gi / 8);
    const int nt = (g.sx / 32) * (g.sy / 8), ntx = g.sx / 32;
    const int count = min(classCount[cls], maxUnits);
    const int2 *list = classList + size_t(cls) * maxUnits;
    const size_t planeXY = size_t(g.sy) * g.sx;
    float *const A = s_lds;  // after the layers: four groups' rows [4][G][HWp]
    for (int u = blockIdx.x; u < count; u += gridDim.x) {
        const int2 unit = list[u];
        const int slot = unit.x, b = unit.y;
        const int k = slot / nt, t = slot % nt;
        const int tx = t % ntx, ty = t / ntx;
        // this block's layers in order, from the slot's list (wave 0, lane = list entry)
        if (w == 0) {
            const int *sl = slotLayers + size_t(slot) * g.numLayers;
            int base = 0;
            for (int j0 = 0; j0 < g.numLayers; j0 += 64) {
                const int e = j0 + lane < g.numLayers ? sl[j0 + lane] : -1;
                // entries after the list's terminator are not written for this beam: only the
                // lanes before the first negative entry hold list entries
                const uint64_t ends = __ballot(e < 0);
                const int endLane = ends ? __builtin_ctzll(ends) : 64;
                const bool in = lane < endLane && (e >> 16) == b + 1;
                const uint64_t m = __ballot(in);
                if (in) s_list[base + __popcll(m & ((uint64_t(1) << lane) - 1))] = (e & 0xff) | (((e >> 8) & 0xff) << 16);
                base += __popcll(m);
                if (ends) break;  // end of the list
            }
            if (lane == 0) s_n = base;
        }
        // sources this thread stages: src = tid + i NT
        size_t pixOff[SPT];
        bool pixIn[SPT];
#pragma unroll
        for (int i = 0; i < SPT; ++i) {
            const int src = tid + i * NT;
            const int gx = tx * 32 + (src & 31), gy = ty * 8 + (src >> 5);
            pixIn[i] = src < 256 && gx < g.rx && gy < g.ry;
            pixOff[i] = size_t(gy) * g.sx + gx;
        }
        __syncthreads();
        const int n = s_n;
        float acc[HW];
#pragma unroll
        for (int c = 0; c < HW; ++c) acc[c] = 0.0f;
        // the pixels (d, r) of list entry li
        auto load = [&](int li, float2 (&px)[SPT]) {
            const float2 *plane = iddSigma + (size_t(s_list[li] & 0xffff) * g.K + k) * planeXY;
#pragma unroll
            for (int i = 0; i < SPT; ++i) px[i] = pixIn[i] ? plane[pixOff[i]] : make_float2(0.0f, PBS_INF);
        };
        // list entry li's table: each source's row [d, e_0 .. e_R, 0 ...] and the source rows' live
        // masks. A wave whose 64 sources are all dead writes zero rows without evaluating the
        // weights (the sweep may still read them, next to another row's live chunk)
        auto stage = [&](int li, const float2 (&px)[SPT]) {
            const int R = s_list[li] >> 16;
#pragma unroll
            for (int i = 0; i < SPT; ++i) {
                const int src = tid + i * NT;
                if (src >= 256) break;  // uniform per wave: NT and 256 are multiples of 64
                const float d = px[i].x, r = px[i].y;
                const uint64_t live = __ballot(d != 0.0f);
                if (lane == 0) {
                    s_live[src >> 5] = uint32_t(live);
                    s_live[(src >> 5) + 1] = uint32_t(live >> 32);
                }
                float row[EW];
                row[0] = d;
                if (live) {
                    gauss_row<EW>(row, r, R);
                } else {
#pragma unroll
                    for (int m = 1; m < EW; ++m) row[m] = 0.0f;
                }
                float *dst = s_lds + (src >> 5) * RP + S::col_off(src & 31);
#pragma unroll
                for (int m = 0; m < EW / 4; ++m)
                    *reinterpret_cast<float4 *>(dst + 4 * m) =
                        make_float4(row[4 * m], row[4 * m + 1], row[4 * m + 2], row[4 * m + 3]);
            }
        };
        // the next layer's pixels are loaded while this one is staged and swept
        float2 pre[SPT];
        if (n > 0) load(0, pre);
        for (int li = 0; li < n; ++li) {
            float2 cur[SPT];
#pragma unroll
            for (int i = 0; i < SPT; ++i) cur[i] = pre[i];
            if (li + 1 < n) load(li + 1, pre);
            stage(li, cur);
            lds_barrier();
            const int R = s_list[li] >> 16;
            // skip the layer when no source row of the wave has a live pixel
            uint32_t wmv = 0;
#pragma unroll
            for (int h = 0; h < 64 / G; ++h) wmv |= s_live[((w * 64) / G + h) % 8];
            const uint32_t wm = __builtin_amdgcn_readfirstlane(wmv) & (half ? 0xffff0000u : 0x0000ffffu);
            if (wm) {
                const int v = q < R + 1 ? q : R + 1;
                const int rp = R > S::RLO ? R : S::RLO;
                if (half)
                    rows_dispatch<S::RLO, RM, EW, 1>(rp, acc, s_lds + grp * RP, v, wm);
                else
                    rows_dispatch<S::RLO, RM, EW, 0>(rp, acc, s_lds + grp * RP, v, wm);
            }
            lds_barrier();  // the table is read: the next layer may overwrite it
        }
        // block pixel (Y, X) = the sum over the groups (h, g), in group order, of row |Y - RM - g|
        // of the group at column X - 16h, where |Y - RM - g| <= RM and 0 <= X - 16h < 16 + 2RM;
        // four groups at a time through A
        // this thread's block pixels (Y, X), X = pixel mod BW, from a thread index that an empty
        // asm makes opaque here: the compiler would otherwise compute the pixels' coordinates and
        // LDS addresses before the unit loop and hold them across every layer (~30 registers:
        // 152 instead of 125 for RM 11, three waves per SIMD instead of four)
        int tb = tid;
        asm volatile("" : "+v"(tb));
        float part[OPT];
#pragma unroll
        for (int j = 0; j < OPT; ++j) part[j] = 0.0f;
#pragma unroll 1
        for (int ph = 0; ph < 4; ++ph) {
            if (gi / 4 == ph && q <= RM) {
                float *a = A + ((gi % 4) * G + q) * HWp;
#pragma unroll
                for (int c = 0; c < HW; ++c) a[c] = acc[c];
            }
            __syncthreads();
            const int dx0 = 16 * (ph >> 1), dy0 = RM + 4 * (ph & 1);  // this phase's half and first row
#pragma unroll
            for (int j = 0; j < OPT; ++j) {
                const int o = tb + j * NT;
                if (o >= S::OUT) break;
                const int hx = o % BW - dx0;  // the column within this phase's half
                if (unsigned(hx) < unsigned(HW)) {
                    const int y0 = o / BW - dy0;
                    float sum = part[j];
                    if (y0 >= 3 && y0 <= RM) {
                        // rows y0 - g4 of the four groups, all within reach: a constant stride
                        const float *col = A + y0 * HWp + hx;
#pragma unroll
                        for (int g4 = 0; g4 < 4; ++g4) sum += col[g4 * (G - 1) * HWp];
                    } else {
#pragma unroll
                        for (int g4 = 0; g4 < 4; ++g4) {
                            const int ady = abs(y0 - g4);
                            if (ady <= RM) sum += A[(g4 * G + ady) * HWp + hx];
                        }
                    }
                    part[j] = sum;
                }
            }
            __syncthreads();
        }
        float *out;
        if (b < 0) {
            out = mainSlots + size_t(slot) * kSlotFloats;
        } else {
            const int qq = gs[b].tailBase + (k - gs[b].tailStart);
            out = (qq < bs->tailPlanes) ? tailSlots + (size_t(qq) * nt + t) * kSlotFloats : nullptr;
        }
        if (out) {
#pragma unroll
            for (int j = 0; j < OPT; ++j) {
                const int o = tid + j * NT;
                if (o >= S::OUT) break;
                out[(o / BW) * kSlotPitch + (o % BW)] = part[j];
            }
        }
    }
}
template <int RM>
struct SuperposeRowsK {
    static constexpr int NT = RowsShape<RM>::NT;
    // four waves per SIMD (128 registers) for RM 7 and 11: the one-beam kernels fit without the
    // bound (115 / 125), the batch's unit queue needs it (137 otherwise; a few spilled dwords of
    // its beam loop cost less than the fourth wave gains); three (168) for RM 15
    static constexpr int WPE = RM <= 11 ? 4 : 3;
    static constexpr bool kUnitQueue = true;  // batched as a queue of the beams' units
    struct Args {
        FanGeom g;
        const float2 *iddSigma;
        const int *tileRad;
        const int *effRad;
        const LayerScalars *ls;
        const GroupScalars *gs;
        const BeamScalars *bs;
        const int *classCount;
        const int2 *classList;
        int cls;
        int maxUnits;
        const int *slotLayers;
        float *mainSlots;
        float *tailSlots;
    };
    static __device__ __forceinline__ void run(const Args &a, const uint3 b, const uint3 d) { superpose_rows_body<RM>(b, d, a.g, a.iddSigma, a.tileRad, a.effRad, a.ls, a.gs, a.bs, a.classCount, a.classList, a.cls, a.maxUnits, a.slotLayers, a.mainSlots, a.tailSlots); }
};


// Floor division by a positive divisor.
__device__ __forceinline__ int floor_div(int a, int d) { return (a >= 0 ? a : a - d + 1) / d; }

// Sum of the item blocks covering the BEV pixels of one 32x8 output tile of a plane, in
// (ty, tx) order. The work-group's output tile (otx, oty) sits on the item tiling shifted by
// the plane's 32-pixel margin, so the candidate item tiles are the same for all its pixels
// (RMmax: the plane's largest block layout radius, a tile's block spans 32 + 2RM columns and
// 8 + 2RM rows around it); each lane then adds the one element of each candidate block that
// covers its pixel.
__device__ __forceinline__ float gather_tile(const FanGeom &g, int otx, int oty, int lx, int ly,
                                             const int *__restrict__ layoutR, const float *__restrict__ slots,
                                             int RMmax) {
    const int ntx = g.sx / 32, nty = g.sy / 8;
    const int xr0 = 32 * otx - 32, yr0 = 8 * oty - 32;  // the output tile's first pixel, item-tile frame
    const int tx0 = max(0, floor_div(xr0 - 31 - RMmax, 32)), tx1 = min(ntx - 1, floor_div(xr0 + 31 + RMmax, 32));
    const int ty0 = max(0, floor_div(yr0 - 7 - RMmax, 8)), ty1 = min(nty - 1, floor_div(yr0 + 7 + RMmax, 8));
    const int xr = xr0 + lx, yr = yr0 + ly;
    float sum = 0.0f;
    for (int ty = ty0; ty <= ty1; ++ty)
        for (int tx = tx0; tx <= tx1; ++tx) {
            const int tt = ty * ntx + tx;
            const int RM = layoutR[tt];
            if (RM < 0) continue;
            const int col = xr - 32 * tx + RM, row = yr - 8 * ty + RM;
            if (col >= 0 && col < 32 + 2 * RM && row >= 0 && row < 8 + 2 * RM)
                sum += slots[size_t(tt) * kSlotFloats + row * kSlotPitch + col];
        }
    return sum;
}

// Main BEV [K][sy+64][sx+64] and tail BEV planes [tailCap][sy+64][sx+64]; a work-group per
// 32x8 output tile of a plane (one lane per pixel).
__device__ __forceinline__ void merge_body(const uint3 blockIdx, const uint3 gridDim, FanGeom g, const BeamScalars *__restrict__ bs,
                                               const int *__restrict__ itemR, const float *__restrict__ mainSlots,
                                               float *__restrict__ bev, const int *__restrict__ tailItemR,
                                               const float *__restrict__ tailSlots, float *__restrict__ tailBev,
                                               const int *__restrict__ planeRM) {
    const int bx = g.sx + 64, by = g.sy + 64;
    const int otx = blockIdx.x, oty = blockIdx.y;
    const int lx = threadIdx.x & 31, ly = threadIdx.x >> 5;
    const int X = 32 * otx + lx, Y = 8 * oty + ly;
    const int z = blockIdx.z;
    const int nt = (g.sx / 32) * (g.sy / 8);
    if (z < g.K) {
        // only the planes [beamFirstInside, mainEndMax) hold main items; the transfer reads the
        // others as 0, so they are not written
        if (z < bs->beamFirstInside || z >= bs->mainEndMax) return;
        float sum = 0.0f;
        const int RMmax = planeRM[z];
        if (RMmax >= 0)
            sum = gather_tile(g, otx, oty, lx, ly, itemR + size_t(z) * nt, mainSlots + size_t(z) * nt * kSlotFloats,
                              RMmax);
        bev[(size_t(z) * by + Y) * bx + X] = sum;
    } else {
        // tail planes, strided over the gridDim.z - K tail block rows
        const int tz = int(gridDim.z) - g.K;
        for (int q = z - g.K; q < bs->tailPlanes; q += tz) {
            const int RMmax = planeRM[g.K + q];
            const float sum = RMmax < 0 ? 0.0f
                                        : gather_tile(g, otx, oty, lx, ly, tailItemR + size_t(q) * nt,
                                                      tailSlots + size_t(q) * nt * kSlotFloats, RMmax);
            tailBev[(size_t(q) * by + Y) * bx + X] = sum;
        }
    }
}
struct MergeK {
    static constexpr int NT = 256;
    struct Args {
        FanGeom g;
        const BeamScalars *bs;
        const int *itemR;
        const float *mainSlots;
        float *bev;
        const int *tailItemR;
        const float *tailSlots;
        float *tailBev;
        const int *planeRM;
    };
    static __device__ __forceinline__ void run(const Args &a, const uint3 b, const uint3 d) { merge_body(b, d, a.g, a.bs, a.itemR, a.mainSlots, a.bev, a.tailItemR, a.tailSlots, a.tailBev, a.planeRM); }
};


// ---------------------------------------------------------------------------------------
// BEV -> dose
// ---------------------------------------------------------------------------------------
// Thread per dose voxel inside the union box (one owner per voxel, no atomics). The
// dose-index -> fan-index map is the reference's factored form
// (transfer_param_struct_div3.cuh:34-44, transfer_param_struct_div3_impl.cu:14-34). Every
// voxel samples the main BEV; a voxel inside a group's box also samples that group's tail.
// The dose a beam adds at voxel (x, y, z), or 0 outside its union box.
__device__ __forceinline__ float transfer_at(const FanGeom &g, const TransferArgs &ta,
                                             const BeamScalars *__restrict__ bs, const GroupScalars *__restrict__ gs,
                                             const float *__restrict__ bev, const float *__restrict__ tailBev, int x,
                                             int y, int z) {
    if (x < bs->boxStart[0] || y < bs->boxStart[1] || z < bs->boxStart[2] || x >= bs->boxEnd[0] ||
        y >= bs->boxEnd[1] || z >= bs->boxEnd[2])
        return 0.0f;
    const int inside = bs->beamFirstInside;
    const int m = bs->numGroups;
    const int bx = g.sx + 64, by = g.sy + 64;
    const float fx = float(x), fy = float(y), fz = float(z);
    float px = fx * ta.coefI[0] + fy * ta.coefJ[0] + ta.coefOffset[0] + fz * ta.inc[0];
    float py = fx * ta.coefI[1] + fy * ta.coefJ[1] + ta.coefOffset[1] + fz * ta.inc[1];
    const float pz = fx * ta.coefI[2] + fy * ta.coefJ[2] + ta.coefOffset[2] + fz * ta.inc[2];
    px *= 1.0f + pz / (ta.normDist[0] - pz);
    py *= 1.0f + pz / (ta.normDist[1] - pz);
    const float fxI = px + ta.globalOffset[0], fyI = py + ta.globalOffset[1];
    const float fzI = pz + (ta.globalOffset[2] - float(inside));
    // prim_transf_kernel.cu:28-40: texel coordinate = fan index + 1/2; negatives skipped
    if (fxI + 0.5f < 0.0f || fyI + 0.5f < 0.0f || fzI + 0.5f < 0.0f) return 0.0f;
    if (fxI > float(bx) || fyI > float(by)) return 0.0f;  // wholly outside the BEV: border value 0
    float v = 0.0f;
    if (fzI <= float(g.K))
        v = trilinear_planes(bev, bx, by, inside, min(bs->mainEndMax, g.K), fxI, fyI, fzI);
    for (int j = 0; j < m; ++j) {
        const GroupScalars &G = gs[j];
        if (G.tailCount <= 0) continue;
        const float zl = fzI - float(G.tailStart);
        if (!(zl > -1.0f && zl < float(G.tailCount))) continue;
        if (x < G.boxStart[0] || x >= G.boxEnd[0] || y < G.boxStart[1] || y >= G.boxEnd[1] || z < G.boxStart[2] ||
            z >= G.boxEnd[2])
            continue;
        v += trilinear_border(tailBev + size_t(G.tailBase) * bx * by, bx, by, G.tailCount, fxI, fyI, zl);
    }
    return v;
}

// One beam: a thread per dose voxel inside its union box, a work-group per tx x ty x tz voxel
// brick (256 voxels), short along the beam's depth axis so its samples stay within a few BEV
// planes.
struct TransferK {
    static constexpr int NT = 256;
    struct Args {
        FanGeom g;
        TransferArgs ta;
        const BeamScalars *bs;
        const GroupScalars *gs;
        const float *bev;
        const float *tailBev;
        float *dose;
        int tx, ty, tz;
    };
    static __device__ __forceinline__ void run(const Args &a, const uint3 b, const uint3 d) {
        (void)d;
        const int lx = int(threadIdx.x) % a.tx, ly = (int(threadIdx.x) / a.tx) % a.ty,
                  lz = int(threadIdx.x) / (a.tx * a.ty);
        const int x = a.bs->boxStart[0] + int(b.x) * a.tx + lx;
        const int y = a.bs->boxStart[1] + int(b.y) * a.ty + ly;
        const int z = a.bs->boxStart[2] + int(b.z) * a.tz + lz;
        if (x >= a.bs->boxEnd[0] || y >= a.bs->boxEnd[1] || z >= a.bs->boxEnd[2]) return;
        const float v = transfer_at(a.g, a.ta, a.bs, a.gs, a.bev, a.tailBev, x, y, z);
        if (v > 0.0f) {
            float *res = a.dose + (size_t(z) * a.g.doseDims[1] + y) * a.g.doseDims[0] + x;
            *res += v;
        }
    }
};

// Every transfer of a batch in one launch: a thread per voxel of the whole dose grid (8x8x4
// bricks) adds the beams' values in list order (one owner per voxel, no atomics), so the sum is
// the one the beams make one after another.
__global__ __launch_bounds__(256) void k_transfer_many(const TransferK::Args *__restrict__ list, int n) {
    float *__restrict__ dose = list[0].dose;
    const int dx = list[0].g.doseDims[0], dy = list[0].g.doseDims[1], dz = list[0].g.doseDims[2];
    const int nbx = (dx + 7) / 8, nby = (dy + 7) / 8;
    const int x = int(blockIdx.x % unsigned(nbx)) * 8 + int(threadIdx.x & 7);
    const int y = int((blockIdx.x / unsigned(nbx)) % unsigned(nby)) * 8 + int((threadIdx.x >> 3) & 7);
    const int z = int(blockIdx.x / unsigned(nbx * nby)) * 4 + int(threadIdx.x >> 6);
    if (x >= dx || y >= dy || z >= dz) return;
    float acc = 0.0f;
    bool any = false;
    float *res = dose + (size_t(z) * dy + y) * dx + x;
    for (int i = 0; i < n; ++i) {
        const TransferK::Args &a = list[i];
        const float v = transfer_at(a.g, a.ta, a.bs, a.gs, a.bev, a.tailBev, x, y, z);
        if (v > 0.0f) {
            if (!any) acc = *res;
            any = true;
            acc += v;
        }
    }
    if (any) *res = acc;
}

// Segment blockIdx.y of the upload, 32-bit words strided over the blocks of blockIdx.x (the
// source is host memory: every word is read once).
// The one-pass trace's partials (row a.n of the upload launch): a wave per step takes the minimum
// of the step's per-work-group minima, and wave 0 of work-group 0 the least entry and greatest
// exit step. Minima and maxima are exact, whatever the order.
__device__ __forceinline__ void trace_partials_body(const uint3 blockIdx, const uint3 gridDim, const UploadArgs &a) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int k = int(blockIdx.x) * 4 + w; k < a.K; k += int(gridDim.x) * 4) {
        // the row as float4s: the pitch is a multiple of 4 and the partials start 256-byte aligned
        const float4 *row = reinterpret_cast<const float4 *>(a.tp.min + size_t(k) * a.tp.pitch);
        float m = 1.0e20f;
        for (int i = lane; i < a.tp.pitch / 4; i += 64) {
            const float4 v = row[i];
            m = fminf(m, fminf(fminf(v.x, v.y), fminf(v.z, v.w)));
        }
        m = wave_min_uniform(m);  // the wave is whole again after the loop
        if (lane == 0) a.weplMin[k] = m;
    }
    if (blockIdx.x == 0 && w == 0) {
        int mn = INT_MAX, mx = INT_MIN;
        for (int i = lane; i < a.tp.n; i += 64) {
            mn = min(mn, a.tp.fi[i]);
            mx = max(mx, a.tp.fo[i]);
        }
        mn = wave_min_uniform_i(mn);
        mx = wave_max_uniform_i(mx);
        if (lane == 0) {
            a.bs->beamFirstInside = mn;
            a.bs->beamFirstOutside = mx;
        }
    }
}

__device__ __forceinline__ void upload_body(const uint3 blockIdx, const uint3 gridDim, const UploadArgs &a) {
    const int sg = blockIdx.y;
    if (sg == a.n && a.tp.min) {
        trace_partials_body(blockIdx, gridDim, a);
        return;
    }
    if (sg >= a.n) return;
    const UploadSeg u = a.seg[sg];
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < u.words; i += gridDim.x * 256u) u.dst[i] = u.src[i];
}
struct UploadK {
    static constexpr int NT = 256;
    struct Args {
        UploadArgs a;
    };
    static __device__ __forceinline__ void run(const Args &a, const uint3 b, const uint3 d) { upload_body(b, d, a.a); }
};

// ---------------------------------------------------------------------------------------
// one-beam and multi-beam launches of a kernel body K
// ---------------------------------------------------------------------------------------
// Waves per SIMD a kernel body is compiled for (its register budget): K::WPE, else no bound.
template <typename K, typename = void>
struct WavesPerEu {
    static constexpr int value = 1;
};
template <typename K>
struct WavesPerEu<K, std::void_t<decltype(K::WPE)>> {
    static constexpr int value = K::WPE;
};

template <typename K>
__global__ __launch_bounds__(K::NT) __attribute__((amdgpu_waves_per_eu(WavesPerEu<K>::value))) void k_one(
    typename K::Args a) {
    K::run(a, uint3{blockIdx.x, blockIdx.y, blockIdx.z}, uint3{gridDim.x, gridDim.y, gridDim.z});
}

// Work-group blockIdx.x of the concatenated grids: its beam is the last whose first work-group
// is not after it (a scalar walk over at most a batch of entries), its index within that
// beam's grid is unflattened x-fastest.
template <typename K>
__global__ __launch_bounds__(K::NT) __attribute__((amdgpu_waves_per_eu(WavesPerEu<K>::value))) void k_many(
    const typename K::Args *__restrict__ args, MultiTable t) {
    const unsigned w = blockIdx.x;
    int b = 0;
    if (t.nb <= 64) {
        // lane i reads the end of beam i: the beam is the number of ends at or before w
        const int lane = int(threadIdx.x & 63);
        const unsigned end = lane < t.nb ? t.base[lane + 1] : 0xffffffffu;
        b = __builtin_amdgcn_readfirstlane(__popcll(__ballot(w >= end)));
    } else {
        while (b + 1 < t.nb && w >= t.base[b + 1]) ++b;
    }
    const uint4 d = t.dims[b];
    const unsigned local = w - t.base[b];
    const uint3 id{local % d.x, (local / d.x) % d.y, local / (d.x * d.y)};
    K::run(args[b], id, uint3{d.x, d.y, d.z});
}

// The radius-class kernels over a batch: a fixed grid walks the beams in order, and each beam's
// units (counted on the device) are spread over the whole grid, the beam's first unit on the
// work-group after the previous beam's last, so the grid stays balanced whatever the counts.
// The class bodies end every unit with a barrier, so a work-group may run them back to back.
template <typename K>
__global__ __launch_bounds__(K::NT) __attribute__((amdgpu_waves_per_eu(WavesPerEu<K>::value))) void k_many_units(
    const typename K::Args *__restrict__ args, int nb) {
    const unsigned G = gridDim.x;
    unsigned start = 0;
    for (int b = 0; b < nb; ++b) {
        const typename K::Args &a = args[b];
        const int count = min(a.classCount[a.cls], a.maxUnits);
        if (count <= 0) continue;
        const unsigned id = (blockIdx.x + G - start % G) % G;
        K::run(a, uint3{id, 0u, 0u}, uint3{G, 1u, 1u});
        start += unsigned(count);
    }
}

}  // namespace pbs

// ---------------------------------------------------------------------------------------
// launch wrappers (host)
// ---------------------------------------------------------------------------------------
namespace pbs {

thread_local Recorder *t_rec = nullptr;

template <typename K>
static void launch_one(const void *a, dim3 grid, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL(k_one<K>, grid, dim3(K::NT), lds, s, *static_cast<const typename K::Args *>(a));
}

template <typename K, typename = void>
struct ClassKernel : std::false_type {};
template <typename K>
struct ClassKernel<K, std::void_t<decltype(K::kUnitQueue)>> : std::true_type {};

// Work-groups of a radius class's batched launch: enough to fill the device several times over.
constexpr unsigned kClassBatchGrid = 2048;

template <typename K>
static void launch_many(const void *a, const MultiTable &t, unsigned blocks, size_t lds, hipStream_t s) {
    if constexpr (ClassKernel<K>::value) {
        const unsigned g = blocks < kClassBatchGrid ? blocks : kClassBatchGrid;
        hipLaunchKernelGGL(k_many_units<K>, dim3(g), dim3(K::NT), lds, s, static_cast<const typename K::Args *>(a),
                           t.nb);
    } else {
        hipLaunchKernelGGL(k_many<K>, dim3(blocks), dim3(K::NT), lds, s, static_cast<const typename K::Args *>(a), t);
    }
}

// The batch's transfers: one owner per voxel of the whole dose grid, t.nb entries in order
// (`blocks` bricks of 8x8x4 voxels, from many_blocks).
template <>
void launch_many<TransferK>(const void *a, const MultiTable &t, unsigned blocks, size_t lds, hipStream_t s) {
    (void)lds;
    hipLaunchKernelGGL(k_transfer_many, dim3(blocks), dim3(256), 0, s, static_cast<const TransferK::Args *>(a), t.nb);
}

// Work-groups of a multi-beam launch of K: 0 means the beams' grids concatenated.
template <typename K>
static unsigned many_blocks(const typename K::Args &) { return 0; }
template <>
unsigned many_blocks<TransferK>(const TransferK::Args &a) {
    const unsigned nx = unsigned(a.g.doseDims[0] + 7) / 8, ny = unsigned(a.g.doseDims[1] + 7) / 8;
    return nx * ny * (unsigned(a.g.doseDims[2] + 3) / 4);
}
// Records that may share a launch also share this (the transfers: their dose grid).
template <typename K>
static const void *many_target(const typename K::Args &) { return nullptr; }
template <>
const void *many_target<TransferK>(const TransferK::Args &a) { return a.dose; }

// Launches kernel K on `grid` now, or records it when a Recorder is installed (key: the current
// stage and the launch's index within it, unless `key` is given).
template <typename K>
static void dispatch(const typename K::Args &a, dim3 grid, size_t lds, hipStream_t s, int key = -1) {
    const bool empty = grid.x == 0 || grid.y == 0 || grid.z == 0;
    if (!t_rec) {
        if (!empty) launch_one<K>(&a, grid, lds, s);
        return;
    }
    // an empty launch still takes its place in the stage, so the keys of later launches stay
    // aligned with the other beams'
    const int k = key >= 0 ? key : t_rec->stage * 64 + t_rec->sub++;
    if (empty) return;
    LaunchRec r;
    r.key = k;
    r.beam = t_rec->beam;
    r.seq = t_rec->seq++;
    r.one = &launch_one<K>;
    r.many = &launch_many<K>;
    r.args.resize(sizeof(typename K::Args));
    std::memcpy(r.args.data(), &a, sizeof(typename K::Args));
    r.argAlign = alignof(typename K::Args);
    r.grid = grid;
    r.lds = lds;
    r.manyBlocks = many_blocks<K>(a);
    r.target = many_target<K>(a);
    r.par = t_rec->par;
    t_rec->recs.push_back(std::move(r));
}

// ---------------------------------------------------------------------------------------
// batch executor
// ---------------------------------------------------------------------------------------
namespace {

struct RecGroup {
    size_t first, count;  // a run of the sorted records
};

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Records sorted by (key, kernel, target, beam, record index); a group is a run of equal key,
// kernel and target. A beam has one record per key (its transfers excepted, which share the
// kernel and go in beam order), so launching the groups in this order keeps every beam's own
// order; records of one key but different kernels (beams of different table sets or branches)
// do not depend on each other.
void plan_groups(const std::vector<LaunchRec> &recs, std::vector<size_t> &order, std::vector<RecGroup> &groups) {
    order.resize(recs.size());
    for (size_t i = 0; i < recs.size(); ++i) order[i] = i;
    const std::less<const void *> before;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        const LaunchRec &x = recs[a], &y = recs[b];
        if (x.key != y.key) return x.key < y.key;
        const void *mx = reinterpret_cast<const void *>(x.many), *my = reinterpret_cast<const void *>(y.many);
        if (mx != my) return before(mx, my);
        if (x.target != y.target) return before(x.target, y.target);
        if (x.beam != y.beam) return x.beam < y.beam;
        return x.seq < y.seq;
    });
    groups.clear();
    for (size_t i = 0; i < order.size(); ++i) {
        const LaunchRec &r = recs[order[i]];
        if (!groups.empty()) {
            const LaunchRec &h = recs[order[groups.back().first]];
            if (h.key == r.key && h.many == r.many && h.target == r.target) {
                ++groups.back().count;
                continue;
            }
        }
        groups.push_back(RecGroup{i, 1});
    }
}

// Table bytes of a group: the Args array, the first work-groups, the grids (256-byte aligned).
size_t group_bytes(size_t argSize, size_t n) {
    return align_up(align_up(argSize * n, 16) + align_up(sizeof(unsigned) * (n + 1), 16) + sizeof(uint4) * n, 256);
}

}  // namespace

size_t records_table_bytes(const std::vector<LaunchRec> &recs) {
    std::vector<size_t> order;
    std::vector<RecGroup> groups;
    plan_groups(recs, order, groups);
    size_t b = 0;
    for (const RecGroup &g : groups) b += group_bytes(recs[order[g.first]].args.size(), g.count);
    return b;
}

size_t stage_records(const std::vector<LaunchRec> &recs, unsigned char *hostTable, std::vector<RecLaunch> &out) {
    std::vector<size_t> order;
    std::vector<RecGroup> groups;
    plan_groups(recs, order, groups);
    out.clear();
    size_t off = 0;
    for (size_t gi = 0; gi < groups.size(); ++gi) {
        const RecGroup &g = groups[gi];
        const LaunchRec &head = recs[order[g.first]];
        const size_t argSize = head.args.size();
        unsigned char *h = hostTable + off;
        const size_t baseOff = align_up(argSize * g.count, 16);
        const size_t dimsOff = baseOff + align_up(sizeof(unsigned) * (g.count + 1), 16);
        unsigned *base = reinterpret_cast<unsigned *>(h + baseOff);
        uint4 *dims = reinterpret_cast<uint4 *>(h + dimsOff);
        unsigned tot = 0;
        size_t lds = 0;
        for (size_t j = 0; j < g.count; ++j) {
            const LaunchRec &r = recs[order[g.first + j]];
            std::memcpy(h + j * argSize, r.args.data(), argSize);
            base[j] = tot;
            dims[j] = uint4{r.grid.x, r.grid.y, r.grid.z, 0u};
            tot += r.grid.x * r.grid.y * r.grid.z;
            lds = std::max(lds, r.lds);
        }
        base[g.count] = tot;
        out.push_back(RecLaunch{head.many, off, baseOff, dimsOff, int(g.count), head.manyBlocks ? head.manyBlocks : tot,
                                lds, head.par});
        off += group_bytes(argSize, g.count);
    }
    return off;
}

hipError_t issue_records(const std::vector<RecLaunch> &launches, unsigned char *devTable, hipStream_t s,
                         const ForkJoin *fj) {
    auto issue = [&](const RecLaunch &l, hipStream_t st) {
        unsigned char *d = devTable + l.off;
        const MultiTable t{reinterpret_cast<const unsigned *>(d + l.baseOff),
                           reinterpret_cast<const uint4 *>(d + l.dimsOff), l.nb};
        l.many(d, t, l.blocks, l.lds, st);
    };
    size_t i = 0;
    while (i < launches.size()) {
        size_t j = i + 1;
        if (launches[i].par)
            while (j < launches.size() && launches[j].par) ++j;
        if (j - i == 1 || !fj) {
            for (size_t q = i; q < j; ++q) issue(launches[q], s);
        } else {
            // fork: launch q of the run goes to stream (q - i) mod (1 + sides); join before the next
            hipError_t e = hipEventRecord(fj->fork, s);
            if (e != hipSuccess) return e;
            const size_t lanes = std::min<size_t>(j - i, 1 + kSideStreams);
            for (size_t k = 1; k < lanes; ++k)
                if ((e = hipStreamWaitEvent(fj->side[k - 1], fj->fork, 0)) != hipSuccess) return e;
            for (size_t q = i; q < j; ++q) {
                const size_t k = (q - i) % lanes;
                issue(launches[q], k ? fj->side[k - 1] : s);
            }
            for (size_t k = 1; k < lanes; ++k) {
                if ((e = hipEventRecord(fj->join[k - 1], fj->side[k - 1])) != hipSuccess) return e;
                if ((e = hipStreamWaitEvent(s, fj->join[k - 1], 0)) != hipSuccess) return e;
            }
        }
        i = j;
    }
    return hipGetLastError();
}

hipError_t launch_upload(const UploadArgs &a, hipStream_t s) {
    const bool reduce = a.tp.min != nullptr;
    if (a.n <= 0 && !reduce) return hipSuccess;
    uint32_t most = 0;
    for (int i = 0; i < a.n; ++i) most = most > a.seg[i].words ? most : a.seg[i].words;
    unsigned bx = std::max(1u, std::min(64u, (most + 255u) / 256u));
    if (reduce) bx = std::max(bx, unsigned(a.K + 3) / 4u);  // a wave per step in the reduction row
    dispatch<UploadK>({a}, dim3(bx, unsigned(a.n) + (reduce ? 1u : 0u)), 0, s);
    return hipGetLastError();
}

// Sub-buffer offsets of the trace scratch in floats, each rounded up to 64 floats (256 B) so
// the float2 and double arrays are aligned whatever nray * K is.
struct TraceScratch {
    size_t hu, sp, shu, last, carry, total;
};
static TraceScratch trace_layout(const FanGeom &g) {
    auto up = [](size_t n) { return (n + 63) / 64 * 64; };
    const size_t nray = size_t(g.rx) * g.ry;
    const size_t S = size_t((g.K + kTraceSeg - 1) / kTraceSeg);
    TraceScratch t;
    t.hu = 0;
    t.sp = up(nray * g.K);            // huPrefix [K][nray]
    t.shu = t.sp + up(2 * nray * S);  // segSp float2 [S][nray]
    t.last = t.shu + up(nray * S);    // segHu [S][nray]
    t.carry = t.last + up(nray * S);  // segLast int [S][nray]
    t.total = t.carry + up(2 * nray * S);  // segCarry double [S][nray]
    return t;
}

// The one-pass trace's partials in the scratch: the per-step minima [K][pitch] (pitch: the
// work-groups rounded up to 4, the pad columns at 1e20), then the entry and exit steps.
static int trace_groups(const FanGeom &g) { return (g.rx * g.ry + kTraceRays - 1) / kTraceRays; }
static int trace_pitch(const FanGeom &g) { return (trace_groups(g) + 3) / 4 * 4; }
static size_t trace_partial_floats(const FanGeom &g) {
    return size_t(g.K) * trace_pitch(g) + 2 * size_t(trace_pitch(g));
}

size_t trace_scratch_floats(const FanGeom &g) { return std::max(trace_layout(g).total, trace_partial_floats(g)); }

template <int CH>
static void launch_trace_one_pass(const FanGeom &g, const float *ct, const LutDev &lut, float *density, float *wepl,
                                  int *fi, int *fo, float *scratch, TracePartials *parts, hipStream_t s) {
    const int groups = trace_groups(g), pitch = trace_pitch(g);
    float *partMin = scratch;
    int *partFi = reinterpret_cast<int *>(scratch + size_t(g.K) * pitch);
    int *partFo = partFi + pitch;
    const int S = (g.K + kTraceSeg - 1) / kTraceSeg;
    const size_t shm = sizeof(float) * (size_t(lut.nDensity + lut.nSp + 3) / 4 * 4) +
                       size_t(S) * kTraceRays * (sizeof(float2) + sizeof(float) + sizeof(int)) + sizeof(int) * kTraceRays;
    dispatch<TraceK<CH>>({g, ct, lut, density, wepl, fi, fo, partMin, pitch, partFi, partFo}, dim3(groups), shm, s);
    *parts = TracePartials{partMin, partFi, partFo, pitch, groups};
}

hipError_t launch_trace(const FanGeom &g, const float *ct, const LutDev &lut, float *density, float *wepl, int *fi,
                        int *fo, float *scratch, float *weplMin, BeamScalars *bs, TracePartials *parts,
                        hipStream_t s) {
    *parts = TracePartials{nullptr, nullptr, nullptr, 0, 0};
    if (g.K <= 2 * kTraceSeg * kTraceThreadsPerRay) {
        launch_trace_one_pass<2>(g, ct, lut, density, wepl, fi, fo, scratch, parts, s);
        return hipGetLastError();
    }
    if (g.K <= kMaxTraceSteps) {
        launch_trace_one_pass<4>(g, ct, lut, density, wepl, fi, fo, scratch, parts, s);
        return hipGetLastError();
    }
    const int n = g.rx * g.ry;
    const int S = (g.K + kTraceSeg - 1) / kTraceSeg;
    const TraceScratch L = trace_layout(g);
    float *huPrefix = scratch + L.hu;
    float2 *segSp = reinterpret_cast<float2 *>(scratch + L.sp);
    float *segHu = scratch + L.shu;
    int *segLast = reinterpret_cast<int *>(scratch + L.last);
    double *segCarry = reinterpret_cast<double *>(scratch + L.carry);
    const size_t shm = sizeof(float) * size_t(lut.nDensity + lut.nSp);
    unsigned *minBits = reinterpret_cast<unsigned *>(weplMin);
    dispatch<TraceSegK>({g, ct, lut, density, wepl, huPrefix, segSp, segHu, segLast, bs}, dim3((n + 255) / 256, S), shm, s);
    // the carry grid also covers the K steps whose minimum it initialises
    dispatch<TraceCarryK>({g, huPrefix, segSp, segHu, segLast, segCarry, fi, fo, minBits, bs}, dim3((std::max(n, g.K) + 255) / 256), 0, s);
    dispatch<TraceFixK>({g, segCarry, wepl, fo, minBits}, dim3((n + 255) / 256, g.K), 0, s);
    return hipGetLastError();
}

hipError_t launch_weq_fill(const FanGeom &g, const WeqArgs &a, float *density, float *wepl, int *fi, int *fo,
                          hipStream_t s) {
    const int n = g.rx * g.ry;
    hipLaunchKernelGGL(k_weq_fill, dim3((n + 255) / 256), dim3(256), 0, s, g, a, density, wepl, fi, fo);
    return hipGetLastError();
}

hipError_t launch_trace_reduce(const FanGeom &g, const float *wepl, const int *fi, const int *fo, float *weplMin,
                               BeamScalars *bs, hipStream_t s) {
    hipLaunchKernelGGL(k_trace_reduce, dim3(g.K + 1), dim3(256), 0, s, g, wepl, fi, fo, weplMin, bs);
    return hipGetLastError();
}

hipError_t launch_beam_setup(const FanGeom &g, const float *weplMin, const LayerHost *lh, BeamScalars *bs,
                             LayerScalars *ls, const float *subspots, int maxSub, float *subEntry, float nucFactor,
                             int *classCount, int *hist, hipStream_t s) {
    dispatch<BeamSetupK>({g, weplMin, lh, bs, ls, subspots, maxSub, subEntry, nucFactor, classCount, hist}, dim3(1), 0, s);
    return hipGetLastError();
}

hipError_t launch_conv(const ConvArgs &a, const float *spot, float *interm, float *rayW, const LayerScalars *ls,
                       const BeamScalars *bs, const float *subspots, const float *subEntry, int numLayers,
                       hipStream_t s) {
    dispatch<ConvXK>({a, spot, interm, ls, bs, subspots, subEntry}, dim3((a.rx + 63) / 64, a.spotY, numLayers * a.nSub), 0, s);
    dispatch<ConvYK>({a, interm, rayW, ls, bs, subspots, subEntry}, dim3((a.rx + 63) / 64, a.ry, numLayers), 0, s);
    return hipGetLastError();
}

hipError_t launch_cpb(const CpbArgs &a, const float *sub, float *cpb, float *rays, int rnx, int rny, float rcx,
                      float rcy, float rrx, float rry, int numLayers, hipStream_t s) {
    hipLaunchKernelGGL(k_subspot_cpb, dim3((a.nx * a.ny + 255) / 256, numLayers), dim3(256), 0, s, a, sub, cpb);
    if (rays)
        hipLaunchKernelGGL(k_cpb_to_rays, dim3((rnx * rny + 255) / 256, numLayers), dim3(256), 0, s, a, cpb, rays, rnx,
                           rny, rcx, rcy, rrx, rry);
    return hipGetLastError();
}

hipError_t launch_fill(const FanGeom &g, const LutDev &lut, const LayerHost *lh, LayerScalars *ls,
                       const BeamScalars *bs, const float *density, const float *wepl, const float *rayW, const int *fi,
                       const int *fo, float2 *iddSigma, int *firstPassive, const FillConsts &fc, const HaloFill &hf,
                       int *tileCode, hipStream_t s) {
    const int nt = (g.sx / 32) * (g.sy / 8);
    const size_t lds = sizeof(float) * (size_t(lut.nRrl) + 2 * size_t(g.K) + 1 + size_t(hf.rayToNuc ? 6 : 2) * lut.nS);
    if (hf.rayToNuc)
        dispatch<FillK<true>>({g, lut, lh, ls, bs, density, wepl, rayW, fi, fo, iddSigma, firstPassive, fc, hf, tileCode}, dim3(nt, g.numLayers), lds, s);
    else
        dispatch<FillK<false>>({g, lut, lh, ls, bs, density, wepl, rayW, fi, fo, iddSigma, firstPassive, fc, hf, tileCode}, dim3(nt, g.numLayers), lds, s);
    return hipGetLastError();
}

hipError_t launch_rescale(const FanGeom &g, LayerScalars *ls, const BeamScalars *bs,
                          const float *rayW, const float2 *unitIdd, const int *unitFp, float2 *iddSigma,
                          int *firstPassive, int *tileCode, hipStream_t s) {
    const int nt = (g.sx / 32) * (g.sy / 8);
    hipLaunchKernelGGL(k_rescale, dim3(nt, g.numLayers), dim3(256), 0, s, g, ls, bs, rayW, unitIdd, unitFp, iddSigma,
                       firstPassive, tileCode);
    return hipGetLastError();
}

hipError_t launch_seed_grid(int numLayers, const LayerScalars *ls, const BeamScalars *bs, LayerScalars *lsOut,
                            BeamScalars *bsOut, int *hist, int *classCount, hipStream_t s) {
    dispatch<SeedGridK>({numLayers, ls, bs, lsOut, bsOut, hist, classCount}, dim3(1), 0, s);
    return hipGetLastError();
}

// The profile table at (channel c, fractional depth d, fractional row r), d and r already clamped
// to the table: linear texture filtering at texel centres with clamped addressing.
__device__ __forceinline__ float profile_at(const ProfileArgs &a, int c, float d, float r) {
    const int d0 = int(d), r0 = int(r);  // d, r >= 0
    const int d1 = min(d0 + 1, a.depthN - 1), r1 = min(r0 + 1, a.rows - 1);
    const float fd = d - float(d0), fr = r - float(r0);
    const float *t0 = a.table + size_t(r0) * a.depthN * a.channels + c;
    const float *t1 = a.table + size_t(r1) * a.depthN * a.channels + c;
    const float a0 = t0[size_t(d0) * a.channels], b0 = t0[size_t(d1) * a.channels];
    const float a1 = t1[size_t(d0) * a.channels], b1 = t1[size_t(d1) * a.channels];
    const float v0 = a0 + (b0 - a0) * fd, v1 = a1 + (b1 - a1) * fd;
    return v0 + (v1 - v0) * fr;
}

// Thread per (ray, layer), steps [beamFirstInside, afterLast): the superposition reads no other
// step of the field.
__global__ __launch_bounds__(256) void k_profile_hooks(FanGeom g, ProfileArgs a, const float *__restrict__ wepl,
                                                       float2 *__restrict__ iddSigma,
                                                       const LayerScalars *__restrict__ ls,
                                                       const BeamScalars *__restrict__ bs) {
    const int l = blockIdx.y;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= g.rx * g.ry) return;
    const int x = r % g.rx, y = r / g.rx;
    const float4 lp = a.layer[l];
    const int k0 = max(bs->beamFirstInside, 0), k1 = min(ls[l].afterLast, g.K);
    const size_t planeS = size_t(g.sx) * g.sy, planeR = size_t(g.rx) * g.ry;
    float2 *f = iddSigma + size_t(l) * g.K * planeS + size_t(y) * g.sx + x;
    const int nG = (a.channels - 1) / 2;
    // distance from the virtual source along this divergent ray, per unit of axial distance
    const float sad = 0.5f * (g.dist[0] + g.dist[1]);
    const float fanX = g.fc[0] + float(x) * g.fd[0], fanY = g.fc[1] + float(y) * g.fd[1];
    const float slant = sqrtf(1.0f + (fanX * fanX) / (g.dist[0] * g.dist[0]) + (fanY * fanY) / (g.dist[1] * g.dist[1]));
    for (int k = k0; k < k1; ++k) {
        float2 v = f[size_t(k) * planeS];
        if (!(v.x > 0.0f)) {
            if (a.hooks & 1) {
                v.y = PBS_INF;
                f[size_t(k) * planeS] = v;
            }
            continue;
        }
        const float w = wepl[size_t(k) * planeR + r];
        const float di = fminf(fmaxf((w - a.depth0) / a.depthStep, 0.0f), float(a.depthN - 1));
        if (a.hooks & 1) {
            float sw = 0.0f, s2 = 0.0f;
            for (int q = 0; q < nG; ++q) {
                const float wq = profile_at(a, q, di, lp.x), sq = profile_at(a, q + nG, di, lp.x);
                sw += wq;
                s2 += wq * sq * sq;
            }
            if (!(sw > 1.0e-8f)) {
                v.y = PBS_INF;
            } else {
                const float fanZ = g.fc[2] + float(k) * g.fd[2];
                const float phy = (sad - fanZ) * slant - (sad - a.beamParaPos);
                const float initR2 = lp.y + 2.0f * lp.z * phy + lp.w * phy * phy;
                const float axis = sqrtf(fmaxf(0.0f, 0.5f * fmaxf(0.0f, initR2 + s2 / sw)));
                const float width = 0.5f * (g.fd[0] * (1.0f - fanZ / g.dist[0]) + g.fd[1] * (1.0f - fanZ / g.dist[1]));
                v.y = 0.5f * width / (1.41421356f * (axis + 0.21f));
            }
        }
        if (a.hooks & 2) v.x *= fmaxf(0.0f, profile_at(a, a.channels - 1, di, lp.x));
        f[size_t(k) * planeS] = v;
    }
}

hipError_t launch_profile_hooks(const FanGeom &g, const ProfileArgs &a, const float *wepl, float2 *iddSigma,
                                const LayerScalars *ls, const BeamScalars *bs, hipStream_t s) {
    const int nray = g.rx * g.ry;
    hipLaunchKernelGGL(k_profile_hooks, dim3(unsigned((nray + 255) / 256), unsigned(g.numLayers)), dim3(256), 0, s, g,
                       a, wepl, iddSigma, ls, bs);
    return hipGetLastError();
}

hipError_t launch_clear_pairs(float2 *v, size_t n, hipStream_t s) {
    const size_t blocks = (n + 255) / 256;
    dispatch<ClearPairsK>({v, n}, dim3(unsigned(blocks < 4096 ? blocks : 4096)), 0, s);
    return hipGetLastError();
}

hipError_t launch_finalize_and_tiles(const FanGeom &g, const float2 *iddSigma, LayerScalars *ls, GroupScalars *gs,
                                     BeamScalars *bs, int *tailPlaneGroup, int *tailPlaneK, int tailCap, int *tileRad,
                                     int *hist, int *effRad, int *itemR, int *tailItemR, int *classCount,
                                     int2 *classList, int maxUnits, int *overflowFlag, int codesFromFill,
                                     int *planeRM, int *slotLayers, hipStream_t s) {
    const int nt = (g.sx / 32) * (g.sy / 8);
    const LayerFinalizeK::Args fin{g, ls, gs, bs, tailPlaneGroup, tailPlaneK, tailCap, planeRM};
    if (codesFromFill) {
        dispatch<FinalizeTilesK>({fin, tileRad, hist, effRad, overflowFlag}, dim3(g.numLayers + 1), 0, s);
    } else {
        dispatch<LayerFinalizeK>(fin, dim3(1), 0, s);
        dispatch<TilesK>({g, iddSigma, ls, bs, tileRad, hist}, dim3((nt + 3) / 4, g.K), 0, s);
        dispatch<BinsK>({g.numLayers, hist, effRad, overflowFlag}, dim3((g.numLayers + 63) / 64), 0, s);
    }
    const size_t itemsLds =
        sizeof(int) * (size_t(g.numLayers) * (3 + kMaxRadDev + 2) + 3 * size_t(std::max(g.numLayers, 1)));
    dispatch<ItemsK>({g, tileRad, effRad, ls, gs, bs, itemR, tailItemR, classCount, classList, maxUnits, planeRM, slotLayers}, dim3((nt + 255) / 256, g.K), itemsLds, s);
    return hipGetLastError();
}

template <int RM>
static void launch_class(const FanGeom &g, const float2 *iddSigma, const int *tileRad, const int *effRad,
                         const LayerScalars *ls, const GroupScalars *gs, const BeamScalars *bs, int *classCount,
                         const int2 *classList, int cls, int maxUnits, const int *slotLayers, float *mainSlots,
                         float *tailSlots, int grid, hipStream_t s) {
    if constexpr (RM <= 15)
        dispatch<SuperposeRowsK<RM>>({g, iddSigma, tileRad, effRad, ls, gs, bs, classCount, classList, cls, maxUnits, slotLayers, mainSlots, tailSlots}, dim3(grid), 0, s);
    else
        dispatch<SuperposeK<RM>>({g, iddSigma, tileRad, effRad, ls, gs, bs, classCount, classList, cls, maxUnits, mainSlots, tailSlots}, dim3(grid), 0, s);
}

hipError_t launch_superpose(const FanGeom &g, const float2 *iddSigma, const int *tileRad, const int *effRad,
                            const LayerScalars *ls, const GroupScalars *gs, const BeamScalars *bs,
                            int *classCount, const int2 *classList, int maxUnits, const int *slotLayers,
                            float *mainSlots, float *tailSlots, int gridCap,
                            const hipStream_t streams[1 + kSideStreams]) {
    const int grid = maxUnits < gridCap ? maxUnits : gridCap;
    if (t_rec) t_rec->par = true;  // the classes write disjoint slots: recorded as independent
#define PBS_CLASS(RMV, C, S)                                                                                       \
    launch_class<RMV>(g, iddSigma, tileRad, effRad, ls, gs, bs, classCount, classList, C, maxUnits, slotLayers,   \
                      mainSlots, tailSlots, grid, streams[S])
    // class 11 on the engine stream, 7 on one side stream, 15 and the rare 24 / 32 on the other:
    // each side stream costs a wait at the fork and another at the join (a third side stream for
    // 24 / 32 cost more than it overlapped)
    PBS_CLASS(11, 1, 0);
    PBS_CLASS(7, 0, 1);
    PBS_CLASS(15, 2, 2);
    PBS_CLASS(24, 3, 2);
    PBS_CLASS(32, 4, 2);
#undef PBS_CLASS
    if (t_rec) t_rec->par = false;
    return hipGetLastError();
}

hipError_t launch_merge(const FanGeom &g, const BeamScalars *bs, const int *itemR, const float *mainSlots, float *bev,
                        const int *tailItemR, const float *tailSlots, float *tailBev, int tailCap, const int *planeRM,
                        hipStream_t s) {
    const int bx = g.sx + 64, by = g.sy + 64;
    // tail planes over tz block rows, each looping over planes tz apart. tz follows the beam's
    // steps, clamped to 64 .. 128: the head beam's 200 tail planes then take two passes (merge
    // 64 -> 56 us against 64 rows), while shallower beams, with fewer tail planes, keep 64 rows
    // (more would only add rows that find no plane)
    const int tz = std::min(tailCap, std::max(64, std::min(g.K, 128)));
    dispatch<MergeK>({g, bs, itemR, mainSlots, bev, tailItemR, tailSlots, tailBev, planeRM}, dim3(bx / 32, by / 8, g.K + tz), 0, s);
    return hipGetLastError();
}

hipError_t launch_transfer(const FanGeom &g, const TransferArgs &ta, const BeamScalars *bs, const GroupScalars *gs,
                           const float *bev, const float *tailBev, float *dose, int extentX, int extentY, int extentZ,
                           int depthAxis, int shape, hipStream_t s) {
    // 256-voxel bricks: (first cross axis, second cross axis, depth axis)
    static const int kShapes[kNumTransferShapes][3] = {{8, 8, 4}, {16, 16, 1}, {16, 8, 2}, {32, 8, 1}, {8, 4, 8}};
    const int *sh = kShapes[shape >= 0 && shape < kNumTransferShapes ? shape : 0];
    int t[3];
    t[depthAxis] = sh[2];
    t[depthAxis == 0 ? 1 : 0] = sh[0];
    t[depthAxis == 2 ? 1 : 2] = sh[1];
    const int e[3] = {extentX, extentY, extentZ};
    dim3 grid;
    grid.x = unsigned((e[0] + t[0] - 1) / t[0]);
    grid.y = unsigned((e[1] + t[1] - 1) / t[1]);
    grid.z = unsigned((e[2] + t[2] - 1) / t[2]);
    dispatch<TransferK>({g, ta, bs, gs, bev, tailBev, dose, t[0], t[1], t[2]}, grid, 0, s, kTransferKey);
    return hipGetLastError();
}

}  // namespace pbs
