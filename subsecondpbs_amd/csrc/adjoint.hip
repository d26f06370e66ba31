// gfx950 kernels of the adjoint of the pencil-beam dose with respect to the spot weights
// (spot-lattice path, halo off). For fixed live masks, tile radii and step ranges the dose is
// linear in the spot weights, dose = T S F C w (C: spot lattice -> ray weights, F: IDD = ray
// weight x unit IDD, S: superposition, T: BEV -> dose transfer); these kernels apply the
// transposes in reverse order to dJ/dDose:
//
//   k_adj_transfer ..... T^T: each dose voxel scatters its gradient into the 8 BEV cells its
//                        forward sample reads (main BEV and the group tails it samples), as
//                        64-bit fixed-point integers: integer addition is associative, so the
//                        sums do not depend on the order of the atomics and the BEV gradient
//                        (and everything after it) is bitwise repeatable; k_adj_absmax fixes
//                        the scale, k_adj_dequant turns the sums back into floats
//   k_adj_superpose .... S^T and F^T per work unit (step, tile, block): the unit's BEV
//                        window staged in LDS, each source pixel gathers its footprint
//                        sum e_|j| e_|i| dJ/dBEV, times IDD / W (one writer per layer, step
//                        and ray)
//   k_adj_reduce ....... dJ/dW per layer and ray: the steps summed in order
//   k_adj_conv_y / _x .. C^T: the separable erf gathers transposed, walking the forward's
//                        (output, input) pairs (convolution.cu:198-262, 477-513)
#include "kernels.h"

namespace pbs {

// Fixed-point scale of the transposed transfer: 2^(40 - e) for max |dJ/dDose| = m 2^e
// (m in [0.5, 1)). Every contribution w g (w <= 1) is then below 2^40 in magnitude, so up to
// 2^22 contributions to one BEV cell stay inside the int64 range, and the quantum is 2^-40 of the
// largest gradient (float keeps 2^-24 of each value).
constexpr int kAdjFixBits = 40;
__device__ __forceinline__ double adj_fix_scale(unsigned maxBits) {
    int e = 0;
    (void)frexpf(__uint_as_float(maxBits), &e);
    return ldexp(1.0, kAdjFixBits - e);
}

// Scatter g into the in-range corners of a trilinear sample (the transpose of the border-0
// trilinear gather of the transfer; planes outside [zlo, zhi) read as 0 there), in fixed point:
// the float weight times g is exact in double, the power-of-two scale is exact, and one
// rounding to an integer follows.
__device__ __forceinline__ void trilinear_border_adj(unsigned long long *__restrict__ v, int nx, int ny, int zlo,
                                                     int zhi, float x, float y, float z, float g, double scale) {
    const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    const int x0 = int(fx), y0 = int(fy), z0 = int(fz);
    const float tx = x - fx, ty = y - fy, tz = z - fz;
    const double gs = double(g) * scale;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int xi = x0 + (q & 1), yi = y0 + ((q >> 1) & 1), zi = z0 + (q >> 2);
        if (unsigned(xi) < unsigned(nx) && unsigned(yi) < unsigned(ny) && zi >= zlo && zi < zhi) {
            const float w = ((q & 1) ? tx : 1.0f - tx) * (((q >> 1) & 1) ? ty : 1.0f - ty) * ((q >> 2) ? tz : 1.0f - tz);
            const long long c = __double2ll_rn(double(w) * gs);
            if (c) atomicAdd(v + (size_t(zi) * ny + yi) * nx + xi, static_cast<unsigned long long>(c));
        }
    }
}

// max |dJ/dDose| as float bits (non-negative floats order as their bit patterns); *maxBits was
// cleared by k_adj_zero_fix.
__global__ __launch_bounds__(256) void k_adj_absmax(const float *__restrict__ gDose, size_t n,
                                                    unsigned *__restrict__ maxBits) {
    float m = 0.0f;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        m = fmaxf(m, fabsf(gDose[i]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0 && m > 0.0f) atomicMax(maxBits, __float_as_uint(m));
}

__global__ void k_adj_zero_fix(unsigned long long *__restrict__ v, size_t n, unsigned *__restrict__ maxBits) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *maxBits = 0u;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        v[i] = 0ull;
}

// The fixed-point sums back to float dJ/dBEV (one rounding).
__global__ void k_adj_dequant(const unsigned long long *__restrict__ q, size_t n, const unsigned *__restrict__ maxBits,
                              float *__restrict__ out) {
    const unsigned mb = *maxBits;
    const double inv = mb ? 1.0 / adj_fix_scale(mb) : 0.0;  // a power of two: exact
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        out[i] = float(double(static_cast<long long>(q[i])) * inv);
}

// T^T: the voxel -> fan map and sample selection of k_transfer (prim_transf_kernel.cu:11-44).
__global__ __launch_bounds__(256) void k_adj_transfer(FanGeom g, TransferArgs ta, const BeamScalars *__restrict__ bs,
                                                      const GroupScalars *__restrict__ gs,
                                                      const float *__restrict__ gDose,
                                                      unsigned long long *__restrict__ gBev,
                                                      unsigned long long *__restrict__ gTailBev,
                                                      const unsigned *__restrict__ maxBits, int tx, int ty, int tz) {
    const int lx = int(threadIdx.x) % tx, ly = (int(threadIdx.x) / tx) % ty, lz = int(threadIdx.x) / (tx * ty);
    const int x = bs->boxStart[0] + int(blockIdx.x) * tx + lx;
    const int y = bs->boxStart[1] + int(blockIdx.y) * ty + ly;
    const int z = bs->boxStart[2] + int(blockIdx.z) * tz + lz;
    if (x >= bs->boxEnd[0] || y >= bs->boxEnd[1] || z >= bs->boxEnd[2]) return;
    const float gv = gDose[(size_t(z) * g.doseDims[1] + y) * g.doseDims[0] + x];
    if (gv == 0.0f) return;
    const double scale = adj_fix_scale(*maxBits);
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
    if (fxI + 0.5f < 0.0f || fyI + 0.5f < 0.0f || fzI + 0.5f < 0.0f) return;
    if (fxI > float(bx) || fyI > float(by)) return;
    if (fzI <= float(g.K))
        trilinear_border_adj(gBev, bx, by, inside, min(bs->mainEndMax, g.K), fxI, fyI, fzI, gv, scale);
    for (int j = 0; j < m; ++j) {
        const GroupScalars &G = gs[j];
        if (G.tailCount <= 0) continue;
        const float zl = fzI - float(G.tailStart);
        if (!(zl > -1.0f && zl < float(G.tailCount))) continue;
        if (x < G.boxStart[0] || x >= G.boxEnd[0] || y < G.boxStart[1] || y >= G.boxEnd[1] || z < G.boxStart[2] ||
            z >= G.boxEnd[2])
            continue;
        const int nz = min(G.tailCount, bs->tailPlanes - G.tailBase);  // planes that exist
        if (nz > 0)
            trilinear_border_adj(gTailBev + size_t(G.tailBase) * bx * by, bx, by, 0, nz, fxI, fyI, zl, gv, scale);
    }
}

// S^T and F^T for the units of one radius class: dJ/dIDD of a source pixel is its footprint
// sum of dJ/dBEV weighted by e_|j| e_|i| (complete_superposition.cu:298-317 transposed); the
// chain to the ray weight multiplies by IDD / W (IDD = W x unit IDD on live rays). Written to
// gI[l][k][ray]; every (layer, step, ray) belongs to one unit.
constexpr int kAdjMaxR = kMaxRadDev + 1;  // radii up to 33 (the primary overflow bin)
__global__ __launch_bounds__(256) void k_adj_superpose(FanGeom g, const float2 *__restrict__ iddSigma,
                                                       const float *__restrict__ rayW,
                                                       const GroupScalars *__restrict__ gs,
                                                       const BeamScalars *__restrict__ bs,
                                                       const int *__restrict__ classCount,
                                                       const int2 *__restrict__ classList, int cls, int RM,
                                                       int maxUnits, const int *__restrict__ slotLayers,
                                                       const float *__restrict__ gBev,
                                                       const float *__restrict__ gTailBev, float *__restrict__ gI) {
    __shared__ float s_win[(8 + 2 * kAdjMaxR) * (32 + 2 * kAdjMaxR)];
    __shared__ float s_e[(kAdjMaxR + 1) * 256];
    __shared__ int s_list[kMaxLayersDev];
    __shared__ int s_n;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int nt = (g.sx / 32) * (g.sy / 8), ntx = g.sx / 32;
    const int bx = g.sx + 64, by = g.sy + 64;
    const int WW = 32 + 2 * RM, WH = 8 + 2 * RM;
    const int count = min(classCount[cls], maxUnits);
    const int2 *list = classList + size_t(cls) * maxUnits;
    const size_t nsup = size_t(g.sx) * g.sy, nray = size_t(g.rx) * g.ry;
    for (int u = blockIdx.x; u < count; u += gridDim.x) {
        const int2 unit = list[u];
        const int slot = unit.x, b = unit.y;
        const int k = slot / nt, t = slot % nt;
        const int tx = t % ntx, ty = t / ntx;
        const float *plane;
        if (b < 0) {
            plane = gBev + size_t(k) * bx * by;
        } else {
            const int q = gs[b].tailBase + (k - gs[b].tailStart);
            plane = q < bs->tailPlanes ? gTailBev + size_t(q) * bx * by : nullptr;
        }
        if (w == 0) {
            const int *sl = slotLayers + size_t(slot) * g.numLayers;
            int base = 0;
            for (int j0 = 0; j0 < g.numLayers; j0 += 64) {
                const int e = j0 + lane < g.numLayers ? sl[j0 + lane] : -1;
                const uint64_t ends = __ballot(e < 0);
                const int endLane = ends ? __builtin_ctzll(ends) : 64;
                const bool in = lane < endLane && (e >> 16) == b + 1;
                const uint64_t mk = __ballot(in);
                if (in) s_list[base + __popcll(mk & ((uint64_t(1) << lane) - 1))] = e;
                base += __popcll(mk);
                if (ends) break;
            }
            if (lane == 0) s_n = plane ? base : 0;
        }
        // the unit's BEV window: block rows / columns of the item block, plane offset 32 - RM
        const int r0 = ty * 8 + 32 - RM, c0 = tx * 32 + 32 - RM;
        for (int i = tid; i < WH * WW; i += 256) {
            const int rr = i / WW, cc = i % WW;
            s_win[i] = plane ? plane[size_t(r0 + rr) * bx + (c0 + cc)] : 0.0f;
        }
        __syncthreads();
        const int n = s_n;
        const int px = tid & 31, py = tid >> 5;
        const int gx = tx * 32 + px, gy = ty * 8 + py;
        const bool in = gx < g.rx && gy < g.ry;
        for (int li = 0; li < n; ++li) {
            const int e = s_list[li];
            const int l = e & 0xff, R = (e >> 8) & 0xff;
            float c = 0.0f;
            if (in && R <= RM) {  // a radius past the class layout is not superposed by the forward either
                const float2 v = iddSigma[(size_t(l) * g.K + k) * nsup + size_t(gy) * g.sx + gx];
                const float W = rayW[size_t(l) * nray + size_t(gy) * g.rx + gx];
                if (v.x != 0.0f && W >= 1e-6f) {
                    const float r = v.y;
                    float en = erff(r * 0.5f), eo = -en;
                    for (int i = 0; i <= R; ++i) {
                        s_e[i * 256 + tid] = 0.5f * (en - eo);
                        eo = en;
                        if (i < R) en = erff(r * (float(i) + 1.5f));
                    }
                    float acc = 0.0f;
                    for (int j = -R; j <= R; ++j) {
                        const float *row = s_win + (py + RM + j) * WW + (px + RM);
                        float rsum = s_e[tid] * row[0];
                        for (int i = 1; i <= R; ++i) rsum += s_e[i * 256 + tid] * (row[i] + row[-i]);
                        acc += s_e[(j < 0 ? -j : j) * 256 + tid] * rsum;
                    }
                    c = acc * v.x / W;
                }
                gI[(size_t(l) * g.K + k) * nray + size_t(gy) * g.rx + gx] = c;
            }
        }
        __syncthreads();
    }
}

// dJ/dW[l][ray] = sum over the layer's superposed steps of gI, in step order.
__global__ __launch_bounds__(256) void k_adj_reduce(FanGeom g, const LayerScalars *__restrict__ ls,
                                                    const BeamScalars *__restrict__ bs, const float *__restrict__ gI,
                                                    float *__restrict__ gW) {
    const size_t nray = size_t(g.rx) * g.ry;
    const size_t r = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int l = blockIdx.y;
    if (r >= nray) return;
    const int k0 = bs->beamFirstInside, k1 = ls[l].passive;
    float acc = 0.0f;
    for (int k = k0; k < k1; ++k) acc += gI[(size_t(l) * g.K + k) * nray + r];
    gW[size_t(l) * nray + r] = acc;
}

// Transposed erf gather-resample: gIn[ci] = sum over the outputs o whose forward gather
// (erf_gather of k_conv_x / k_conv_y) includes ci, of tap(o, ci) gOut[o]. The forward takes
// ci from ceil((o - reach - off) / delta) while dist = ci delta + off - o < reach; the same
// predicate is evaluated here for every candidate o, so the pair sets are identical.
__device__ __forceinline__ float erf_gather_adj(const float *__restrict__ gOut, int outN, int outStride, int ci,
                                                float sigma, float pixelSp, float off, float delta) {
    const float sEff = sigma / pixelSp;
    const float rS = rsqrtf(2.0f) / sEff;
    const float reach = 3.0f * sEff + 0.5f;
    const float centre = float(ci) * delta + off;
    const int o0 = max(0, int(floorf(centre - reach - delta)) - 1);
    const int o1 = min(outN - 1, int(ceilf(centre + reach)) + 1);
    float res = 0.0f;
    for (int o = o0; o <= o1; ++o) {
        const int cs = int(ceilf((float(o) - reach - off) / delta));
        const float dist = float(ci) * delta + off - float(o);
        if (ci >= cs && dist < reach)
            res += 0.5f * (erff((dist + 0.5f) * rS) - erff((dist - 0.5f) * rS)) * gOut[size_t(o) * outStride];
    }
    return res;
}

// C^T, y pass: gInterm[(l*nSub + s)][cj][x] from gW[l][o][x] (grid (ceil(rx/64), spotY, L*nSub))
__global__ __launch_bounds__(64) void k_adj_conv_y(ConvArgs a, const float *__restrict__ gW,
                                                   float *__restrict__ gInterm, const LayerScalars *__restrict__ ls,
                                                   const BeamScalars *__restrict__ bs,
                                                   const float *__restrict__ subspots,
                                                   const float *__restrict__ subEntry) {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int cj = blockIdx.y;
    const int l = blockIdx.z / a.nSub, s = blockIdx.z % a.nSub;
    if (x >= a.rx) return;
    float sigma = ls[l].entrySigY, off = a.offY;
    bool valid = true;
    if (a.subspotMode) {
        const float *ss = subspots + (size_t(l) * a.nSub + s) * 5;
        valid = ss[2] > 1.0e-8f && ss[3] > 1.0e-8f && ss[4] > 1.0e-8f;
        sigma = subEntry[(size_t(l) * a.nSub + s) * 2 + 1];
        off = (a.spotOffY + ss[1] / a.lenToMm - a.cpbCornerY) / a.cpbResY;
    }
    float r = 0.0f;
    if (valid)
        r = erf_gather_adj(gW + size_t(l) * a.ry * a.rx + x, a.ry, a.rx, cj, sigma, a.cpbResY * bs->pxMultY, off,
                           a.deltaY);
    gInterm[(size_t(blockIdx.z) * a.spotY + cj) * a.rx + x] = r;
}

// C^T, x pass: gSpot[l][y][ci] = sum over subspots of scale x the transposed x gather
// (grid (ceil(spotX/64), spotY, L))
__global__ __launch_bounds__(64) void k_adj_conv_x(ConvArgs a, const float *__restrict__ gInterm,
                                                   float *__restrict__ gSpot, const LayerScalars *__restrict__ ls,
                                                   const BeamScalars *__restrict__ bs,
                                                   const float *__restrict__ subspots,
                                                   const float *__restrict__ subEntry) {
    const int ci = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y, l = blockIdx.z;
    if (ci >= a.spotX) return;
    float acc = 0.0f;
    for (int s = 0; s < a.nSub; ++s) {
        float sigma = ls[l].entrySigX, off = a.offX, scale = 1.0f;
        if (a.subspotMode) {
            const float *ss = subspots + (size_t(l) * a.nSub + s) * 5;
            scale = ss[2];
            sigma = subEntry[(size_t(l) * a.nSub + s) * 2];
            off = (a.spotOffX + ss[0] / a.lenToMm - a.cpbCornerX) / a.cpbResX;
            if (!(scale > 1.0e-8f) || !(ss[3] > 1.0e-8f) || !(ss[4] > 1.0e-8f)) continue;
        }
        const float *row = gInterm + ((size_t(l) * a.nSub + s) * a.spotY + y) * a.rx;
        acc += scale * erf_gather_adj(row, a.rx, 1, ci, sigma, a.cpbResX * bs->pxMultX, off, a.deltaX);
    }
    gSpot[(size_t(l) * a.spotY + y) * a.spotX + ci] = acc;
}

__global__ void k_adj_zero(float *__restrict__ v, size_t n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        v[i] = 0.0f;
}

// ---------------------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------------------
hipError_t launch_adj_zero(float *v, size_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_adj_zero, dim3(2048), dim3(256), 0, s, v, n);
    return hipGetLastError();
}

hipError_t launch_adj_transfer(const FanGeom &g, const TransferArgs &ta, const BeamScalars *bs,
                               const GroupScalars *gs, const float *gDose, size_t doseN, float *gBev,
                               float *gTailBev, size_t mainN, size_t tailN, unsigned long long *fix,
                               unsigned *maxBits, int extentX, int extentY, int extentZ, int depthAxis,
                               hipStream_t s) {
    // fixed-point accumulators [main planes | tail planes], cleared with the scale's maximum
    const size_t nfix = mainN + tailN;
    hipLaunchKernelGGL(k_adj_zero_fix, dim3(2048), dim3(256), 0, s, fix, nfix, maxBits);
    hipLaunchKernelGGL(k_adj_absmax, dim3(1024), dim3(256), 0, s, gDose, doseN, maxBits);
    int t[3] = {8, 8, 8};
    t[depthAxis] = 4;
    const int e[3] = {extentX, extentY, extentZ};
    dim3 grid;
    grid.x = unsigned((e[0] + t[0] - 1) / t[0]);
    grid.y = unsigned((e[1] + t[1] - 1) / t[1]);
    grid.z = unsigned((e[2] + t[2] - 1) / t[2]);
    hipLaunchKernelGGL(k_adj_transfer, grid, dim3(256), 0, s, g, ta, bs, gs, gDose, fix, fix + mainN, maxBits, t[0],
                       t[1], t[2]);
    hipLaunchKernelGGL(k_adj_dequant, dim3(2048), dim3(256), 0, s, fix, mainN, maxBits, gBev);
    if (tailN) hipLaunchKernelGGL(k_adj_dequant, dim3(2048), dim3(256), 0, s, fix + mainN, tailN, maxBits, gTailBev);
    return hipGetLastError();
}

hipError_t launch_adj_superpose(const FanGeom &g, const float2 *iddSigma, const float *rayW, const GroupScalars *gs,
                                const BeamScalars *bs, const int *classCount, const int2 *classList, int maxUnits,
                                const int *slotLayers, const float *gBev, const float *gTailBev, float *gI,
                                hipStream_t s) {
    constexpr int rm[kNumClasses] = {7, 11, 15, 24, 32};
    const int grid = maxUnits < 4096 ? maxUnits : 4096;
    for (int c = 0; c < kNumClasses; ++c)
        hipLaunchKernelGGL(k_adj_superpose, dim3(grid), dim3(256), 0, s, g, iddSigma, rayW, gs, bs, classCount,
                           classList, c, rm[c], maxUnits, slotLayers, gBev, gTailBev, gI);
    return hipGetLastError();
}

hipError_t launch_adj_reduce(const FanGeom &g, const LayerScalars *ls, const BeamScalars *bs, const float *gI,
                             float *gW, hipStream_t s) {
    const int nray = g.rx * g.ry;
    hipLaunchKernelGGL(k_adj_reduce, dim3((nray + 255) / 256, g.numLayers), dim3(256), 0, s, g, ls, bs, gI, gW);
    return hipGetLastError();
}

hipError_t launch_adj_conv(const ConvArgs &a, const float *gW, float *gInterm, float *gSpot, const LayerScalars *ls,
                           const BeamScalars *bs, const float *subspots, const float *subEntry, int numLayers,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_adj_conv_y, dim3((a.rx + 63) / 64, a.spotY, numLayers * a.nSub), dim3(64), 0, s, a, gW,
                       gInterm, ls, bs, subspots, subEntry);
    hipLaunchKernelGGL(k_adj_conv_x, dim3((a.spotX + 63) / 64, a.spotY, numLayers), dim3(64), 0, s, a, gInterm, gSpot,
                       ls, bs, subspots, subEntry);
    return hipGetLastError();
}

}  // namespace pbs
