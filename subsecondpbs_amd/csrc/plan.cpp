// Host-side beam planning (see plan.h). Float operation order follows the reference host
// code so that every integer it derives matches.
#include "plan.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>

namespace pbs {

F3 f3(float x, float y, float z) { return F3{x, y, z}; }
static F3 add(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
static F3 sub(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
static F3 scale(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
static float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static F3 cross(F3 a, F3 b) { return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static F3 normalize_safe(F3 v) {  // wrapper:2886-2890
    const float len = std::sqrt(dot(v, v));
    if (len > 1e-8f) { v.x /= len; v.y /= len; v.z /= len; }
    return v;
}
static F3 mul(const M3 &m, F3 a) { return f3(dot(m.r0, a), dot(m.r1, a), dot(m.r2, a)); }
static M3 transpose(const M3 &m) {
    return M3{f3(m.r0.x, m.r1.x, m.r2.x), f3(m.r0.y, m.r1.y, m.r2.y), f3(m.r0.z, m.r1.z, m.r2.z)};
}
// Matrix product as rows of A dotted with rows of B^T (matrix_3x3.h:66-75).
static M3 mul(const M3 &a, const M3 &b) {
    const M3 t = transpose(b);
    return M3{f3(dot(a.r0, t.r0), dot(a.r0, t.r1), dot(a.r0, t.r2)),
              f3(dot(a.r1, t.r0), dot(a.r1, t.r1), dot(a.r1, t.r2)),
              f3(dot(a.r2, t.r0), dot(a.r2, t.r1), dot(a.r2, t.r2))};
}
static M3 diag(float a, float b, float c) { return M3{f3(a, 0, 0), f3(0, b, 0), f3(0, 0, c)}; }
// Adjugate times 1/det, the latter applied as a scalar matrix (matrix_3x3.h:80-93).
static M3 inverse(const M3 &m) {
    const F3 a = m.r0, b = m.r1, c = m.r2;
    const float det = a.x * (b.y * c.z - b.z * c.y) - a.y * (b.x * c.z - b.z * c.x) + a.z * (b.x * c.y - b.y * c.x);
    const float s = 1.0f / det;
    const M3 adj{f3(b.y * c.z - b.z * c.y, a.z * c.y - a.y * c.z, a.y * b.z - a.z * b.y),
                 f3(b.z * c.x - b.x * c.z, a.x * c.z - a.z * c.x, a.z * b.x - a.x * b.z),
                 f3(b.x * c.y - b.y * c.x, a.y * c.x - a.x * c.y, a.x * b.y - a.y * b.x)};
    return mul(adj, diag(s, s, s));
}
F3 apply(const Aff &t, F3 a) { return add(mul(t.m, a), t.v); }
Aff inverse(const Aff &t) {
    Aff r;
    r.m = inverse(t.m);
    r.v = mul(r.m, scale(t.v, -1.0f));
    return r;
}
Aff concat(const Aff &first, const Aff &second) {
    Aff r;
    r.m = mul(second.m, first.m);
    r.v = add(mul(second.m, first.v), second.v);
    return r;
}

// std::lround for the per-spot lattice cells without the library call: int(v) truncates toward
// zero and v - trunc(v) is exact below 2^23 (above it v is an integer and the remainder is 0), so
// the half cases go away from zero as lround's do. Values beyond 1e9 (and NaN) take lround.
static inline int round_half_away(float v) {
    if (!(std::fabs(v) < 1.0e9f)) return int(std::lround(v));
    int i = int(v);
    const float r = v - float(i);
    if (r >= 0.5f) ++i; else if (r <= -0.5f) --i;
    return i;
}

static int last_smaller_or_eq(const float *v, int n, float value) {
    int upper = n - 1, lower = 0;
    if (v[n - 1] <= value) return upper;
    if (v[0] > value) return -1;
    while (upper - lower > 1) {
        const int pivot = (upper + lower) / 2;
        if (v[pivot] <= value) lower = pivot; else upper = pivot;
    }
    return lower;
}

float find_decimal_ordered(const float *v, int n, float value) {
    if (value >= v[n - 1]) return float(n - 1);
    if (value < v[0]) return 0.0f;
    const unsigned f = unsigned(last_smaller_or_eq(v, n, value));
    const float corr = (value - v[f]) / (v[f + 1] - v[f]);
    return float(f) + corr;
}

// vectorInterpolate (vector_interpolate.h:19-30); also the wrapper's inline lerp (4016-4052).
static float interp_at(const float *v, int n, float idx) {
    if (idx <= 0.0f) return v[0];
    if (idx >= float(n - 1)) return v[n - 1];
    float ip;
    const float dec = std::modf(idx, &ip);
    const unsigned f = unsigned(ip);
    return v[f] + (v[f + 1] - v[f]) * dec;
}

float layer_energy_index(const pbs_energy_struct &e, float energy) {
    const int n = e.nEnergies;
    if (n <= 1) return 0.0f;
    const bool asc = e.energiesPerU[0] < e.energiesPerU[n - 1];
    const int lo = asc ? 0 : n - 1, hi = asc ? n - 1 : 0;
    if (energy <= e.energiesPerU[lo]) return float(lo);
    if (energy >= e.energiesPerU[hi]) return float(hi);
    int fl = 0;
    float e0 = e.energiesPerU[0], e1 = e.energiesPerU[0];
    for (int i = 0; i < n - 1; ++i) {
        const float a = e.energiesPerU[i], b = e.energiesPerU[i + 1];
        const float mn = a < b ? a : b, mx = a < b ? b : a;
        if (energy >= mn && energy <= mx) { fl = i; e0 = a; e1 = b; break; }
    }
    const float den = e1 - e0;
    float corr = den != 0.0f ? (energy - e0) / den : 0.0f;
    if (!std::isfinite(corr)) corr = 0.0f;
    corr = std::min(1.0f, std::max(0.0f, corr));
    return float(fl) + corr;
}

int detect_ct_type(const float *ct, size_t n) {
    // sampleCtStats + autoDetectCtType (wrapper:2307-2335): every (n/1e5)-th voxel
    const size_t stride = n > 100000 ? n / 100000 : 1;
    float mn = ct[0], mx = ct[0];
    for (size_t i = 0; i < n; i += stride) { mn = std::min(mn, ct[i]); mx = std::max(mx, ct[i]); }
    if (mn < -200.0f) return 0;
    if (mx <= 20.0f && mn >= 0.0f) return 2;
    return 1;
}

static float fractional_index_monotonic(const float *v, int n, float value) {
    // findFractionalIndexMonotonic (wrapper:2337-2375): clamp, bisection, linear fraction
    if (n <= 1) return 0.0f;
    const bool inc = v[n - 1] >= v[0];
    if (inc) {
        if (value <= v[0]) return 0.0f;
        if (value >= v[n - 1]) return float(n - 1);
    } else {
        if (value >= v[0]) return 0.0f;
        if (value <= v[n - 1]) return float(n - 1);
    }
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (inc ? (v[mid] <= value) : (v[mid] >= value)) lo = mid; else hi = mid;
    }
    const float d = v[lo + 1] - v[lo];
    float t = d != 0.0f ? (value - v[lo]) / d : 0.0f;
    t = std::min(std::max(t, 0.0f), 1.0f);
    return float(lo) + t;
}

bool ct_to_hu_plus_1000(const float *in, size_t n, int type, const float *lut, int lutN, float scale, float *out) {
    if (type == 0) {
        for (size_t i = 0; i < n; ++i) out[i] = in[i] + 1000.0f;
        return true;
    }
    if (!lut || lutN <= 0 || scale == 0.0f) return false;
    for (size_t i = 0; i < n; ++i) out[i] = fractional_index_monotonic(lut, lutN, in[i]) / scale;
    return true;
}

PlanEnv read_plan_env() {
    PlanEnv env;
    auto lower = [](const char *v) {
        std::string s(v);
        for (char &c : s) c = char(std::tolower(static_cast<unsigned char>(c)));
        return s;
    };
    if (const char *v = std::getenv("RTD_CT_INPUT_TYPE")) {
        const std::string s = lower(v);
        if (s == "hu") env.ctInputType = 0;
        else if (s == "hu+1000" || s == "huplus1000" || s == "hup1000") env.ctInputType = 1;
        else if (s == "density" || s == "rho") env.ctInputType = 2;
        else if (s == "spr" || s == "stoppingpower" || s == "stopping_power") env.ctInputType = 3;
    }
    const char *scale = std::getenv("RTD_ENERGY_DEPTH_SCALE");
    const char *unit = std::getenv("RTD_ENERGY_DEPTH_UNIT");
    if (scale && scale[0] != '\0') {
        const float f = std::max(0.0f, float(std::atof(scale)));
        env.energyDepthMode = f == 0.0f ? 2 : 1;
        env.energyDepthToMm = f;
    } else if (unit && unit[0] != '\0') {
        const std::string u = lower(unit);
        env.energyDepthMode = (u == "cm" || u == "mm") ? 1 : 2;
        env.energyDepthToMm = u == "cm" ? 10.0f : 1.0f;
    }
    if (const char *v = std::getenv("RTD_CPB_COVERAGE_CM")) {
        const float f = float(std::atof(v));
        if (f > 0.0f && std::isfinite(f)) env.cpbCoverage = f;
    }
    return env;
}

void batch_bins(const int *histIn, int *batched, int *effRad) {
    int h[kMaxRad + 2];
    std::memcpy(h, histIn, sizeof(h));
    h[kMaxRad] += h[kMaxRad + 1];  // primary overflow is clamped into the largest radius
    h[kMaxRad + 1] = 0;
    int layerMax = 0;
    for (int i = 0; i < kMaxRad + 2; ++i) if (h[i] > 0) layerMax = i;
    for (int i = 0; i <= kMaxRad; ++i) batched[i] = 0;
    for (int i = 0; i < kMaxRad + 2; ++i) effRad[i] = i;
    effRad[kMaxRad + 1] = kMaxRad;
    batched[0] = h[0];
    int rec = layerMax;
    for (int r = layerMax; r > 0; --r) {
        batched[rec] += h[r];
        effRad[r] = rec;
        if (batched[rec] >= kMinTilesInBatch) rec = r - 1;
    }
}

// buildRawSpotLattice, physical-position mode (wrapper:1259-1473).
static bool place_spots(const pbs_beam_settings &b, BeamPlan &p);

// The spot lattice of the beam (buildRawSpotLattice, wrapper:1366-1460): its axes, then (unless
// `defer`) the placement of every spot on them.
static bool build_lattice(const pbs_beam_settings &b, BeamPlan &p, bool defer) {
    if (!b.layerSpotCounts || !b.spotPositions || !b.spotWeights || b.numSpots <= 0) return false;
    int total = 0;
    for (int l = 0; l < b.numLayers; ++l) total += b.layerSpotCounts[l];
    if (total <= 0 || total != b.numSpots) return false;
    if (b.spotPositionsAreIndices) {
        // index-coded spots (buildRawSpotLattice index mode, wrapper:1295-1364): the lattice is
        // the WEQ ray lattice of the header; each spot is rasterised bilinearly at its texel
        // coordinate - 1/2 (rasterizeContinuousSpotToDensePlane, wrapper:931-968)
        if (!b.waterEquivalence || b.waterEquivalenceCount < 9) return false;
        const float *h = b.waterEquivalence;
        const int nx = std::max(1, int(std::lround(h[8]))), ny = std::max(1, int(std::lround(h[5])));
        std::vector<float> dense(size_t(nx) * ny * b.numLayers, 0.0f);
        int off = 0;
        for (int l = 0; l < b.numLayers; ++l) {
            const int cnt = b.layerSpotCounts[l];
            if (cnt <= 0) return false;
            for (int i = 0; i < cnt; ++i) {
                const int s = off + i;
                const float w = b.spotWeights[s];
                if (!std::isfinite(w) || w < 0.0f) return false;
                if (w == 0.0f) continue;
                const float lx = b.spotPositions[2 * s] - 0.5f, ly = b.spotPositions[2 * s + 1] - 0.5f;
                const int ix0 = int(std::floor(lx)), iy0 = int(std::floor(ly));
                const float tx = lx - float(ix0), ty = ly - float(iy0);
                bool wrote = false;
                auto put = [&](int ix, int iy, float frac) {
                    if (!(frac > 0.0f) || ix < 0 || ix >= nx || iy < 0 || iy >= ny) return;
                    dense[(size_t(l) * ny + iy) * nx + ix] += w * frac;
                    wrote = true;
                };
                put(ix0, iy0, (1.0f - tx) * (1.0f - ty));
                put(ix0 + 1, iy0, tx * (1.0f - ty));
                put(ix0, iy0 + 1, (1.0f - tx) * ty);
                put(ix0 + 1, iy0 + 1, tx * ty);
                if (!wrote) return false;
            }
            off += cnt;
        }
        p.spotCell.assign(size_t(total), -1);  // a spot spreads over up to four cells
        p.spotDims[0] = nx; p.spotDims[1] = ny;
        p.spotOffset[0] = h[6]; p.spotOffset[1] = h[3];
        p.spotDelta[0] = h[7]; p.spotDelta[1] = h[4];
        p.denseWeights.swap(dense);
        return true;
    }
    std::vector<float> xs(total), ys(total);
    for (int i = 0; i < total; ++i) { xs[i] = b.spotPositions[2 * i]; ys[i] = b.spotPositions[2 * i + 1]; }
    // sorted values with neighbours within 1e-4 of the last kept one merged (the unique-x /
    // unique-y step inference, wrapper:1366-1411). Exact repeats (a lattice repeats its positions
    // in every layer) are dropped first through a small hash set, then only the distinct values
    // are sorted: removing exact repeats changes neither the sorted order nor which value heads
    // a merged run
    auto unique_sorted = [](const std::vector<float> &v) {
        size_t cap = 64;
        int capBits = 6;
        while (cap < 2 * v.size()) { cap <<= 1; ++capBits; }
        std::vector<uint32_t> slot(cap, 0u);
        std::vector<char> used(cap, 0);
        std::vector<float> d;
        for (float x : v) {
            const float f = x + 0.0f;  // -0 and +0 are one value
            uint32_t bits;
            std::memcpy(&bits, &f, 4);
            // Fibonacci hashing: the product's high bits (lattice coordinates such as 2.5 or 5.0
            // have all-zero low mantissa bits, so the low bits of the product would all collide)
            size_t h = size_t((bits * 2654435761u) >> (32 - capBits));
            while (used[h] && slot[h] != bits) h = (h + 1) & (cap - 1);
            if (!used[h]) {
                used[h] = 1;
                slot[h] = bits;
                d.push_back(f);
            }
        }
        std::sort(d.begin(), d.end());
        std::vector<float> u;
        for (float x : d) if (u.empty() || !(std::fabs(u.back() - x) <= 1e-4f)) u.push_back(x);
        return u;
    };
    const std::vector<float> ux = unique_sorted(xs), uy = unique_sorted(ys);
    auto step_of = [](const std::vector<float> &u) {
        float s = 1.0e20f;
        for (size_t i = 1; i < u.size(); ++i) {
            const float d = u[i] - u[i - 1];
            if (d > 1.0e-4f && d < s) s = d;
        }
        return s == 1.0e20f ? 1.0f : s;
    };
    const float dx = step_of(ux), dy = step_of(uy), ox = ux.front(), oy = uy.front();
    const int nx = std::max(1, int(std::lround((ux.back() - ox) / dx)) + 1);
    const int ny = std::max(1, int(std::lround((uy.back() - oy) / dy)) + 1);
    for (int l = 0; l < b.numLayers; ++l)
        if (b.layerSpotCounts[l] <= 0) return false;
    p.spotDims[0] = nx; p.spotDims[1] = ny;
    p.spotOffset[0] = ox; p.spotOffset[1] = oy;
    p.spotDelta[0] = dx; p.spotDelta[1] = dy;
    if (defer) {
        p.placementPending = true;
        return true;
    }
    if (place_spots(b, p)) return true;
    p.spotDims[0] = p.spotDims[1] = 0;  // not a lattice beam: no lattice in the plan
    p.spotOffset[0] = p.spotOffset[1] = 0.0f;
    p.spotDelta[0] = p.spotDelta[1] = 0.0f;
    return false;
}

// Every spot of a physical-position lattice onto the axes in p: its cell, and its weight added
// to the dense lattice; false when a weighted spot is off the lattice.
static bool place_spots(const pbs_beam_settings &b, BeamPlan &p) {
    p.placementPending = false;
    const int total = b.numSpots;
    const int nx = p.spotDims[0], ny = p.spotDims[1];
    const float ox = p.spotOffset[0], oy = p.spotOffset[1], dx = p.spotDelta[0], dy = p.spotDelta[1];
    std::vector<float> dense(size_t(nx) * ny * b.numLayers, 0.0f);
    std::vector<int> cells(size_t(total), -1);
    int off = 0;
    for (int l = 0; l < b.numLayers; ++l) {
        const int cnt = b.layerSpotCounts[l];
        if (cnt <= 0) return false;
        for (int i = 0; i < cnt; ++i) {
            const int s = off + i;
            const float x = b.spotPositions[2 * s], y = b.spotPositions[2 * s + 1], w = b.spotWeights[s];
            if (!std::isfinite(w) || w < 0.0f) return false;
            const float fx = dx != 0.0f ? (x - ox) / dx : 0.0f;
            const float fy = dy != 0.0f ? (y - oy) / dy : 0.0f;
            const int ix = round_half_away(fx), iy = round_half_away(fy);
            const bool inLattice = ix >= 0 && ix < nx && iy >= 0 && iy < ny;
            if (inLattice) cells[s] = int((size_t(l) * ny + iy) * nx + ix);
            if (w == 0.0f) continue;
            if (!inLattice) return false;
            if (!(std::fabs(ox + ix * dx - x) <= 2e-2f) || !(std::fabs(oy + iy * dy - y) <= 2e-2f)) return false;
            dense[(size_t(l) * ny + iy) * nx + ix] += w;
        }
        off += cnt;
    }
    p.spotCell.swap(cells);
    p.denseWeights.swap(dense);
    return true;
}

bool place_plan_spots(const pbs_beam_settings &b, BeamPlan &p) { return !p.placementPending || place_spots(b, p); }

// alignGridToReferencePhase (wrapper:970-997).
static bool align_grid(float minC, float maxC, float ref, float step, int multiple, float &origin, int &count) {
    if (!(step > 0.0f) || !std::isfinite(minC) || !std::isfinite(maxC) || !std::isfinite(ref)) return false;
    if (maxC < minC) std::swap(minC, maxC);
    const float eps = 1.0e-4f;
    const int fs = int(std::ceil((minC - ref) / step - eps));
    const int ls = int(std::floor((maxC - ref) / step + eps));
    if (ls < fs) return false;
    origin = ref + float(fs) * step;
    count = std::max(1, ls - fs + 1);
    if (multiple > 1) count = ((count + multiple - 1) / multiple) * multiple;
    return true;
}

// estimateVirtualSourceDistancesFromSpots, physical positions (wrapper:1847-1921).
static bool infer_source_dist(const pbs_beam_settings &b, F3 bmX, F3 bmY, F3 bmZ, float sad, float &sx, float &sy) {
    if (!(sad > 1.0e-6f) || !b.spotBeamDirections || !b.spotPositions || b.numSpots <= 0) return false;
    std::vector<float> ax, ay;
    const bool idx = b.spotPositionsAreIndices != 0;
    if (idx && (!b.waterEquivalence || b.waterEquivalenceCount < 9)) return false;
    for (int i = 0; i < b.numSpots; ++i) {
        float px = b.spotPositions[2 * i], py = b.spotPositions[2 * i + 1];
        if (idx) {  // texel coordinate -> reference-plane coordinate (wrapper:1881-1888)
            const float *h = b.waterEquivalence;
            px = h[6] + (px - 0.5f) * h[7];
            py = h[3] + (py - 0.5f) * h[4];
        }
        F3 d = f3(b.spotBeamDirections[3 * i], b.spotBeamDirections[3 * i + 1], b.spotBeamDirections[3 * i + 2]);
        const float len = std::sqrt(dot(d, d));
        if (!(len > 1.0e-6f)) continue;
        d.x /= len; d.y /= len; d.z /= len;
        const float a = dot(d, bmZ);
        if (!(std::fabs(a) > 1.0e-6f)) continue;
        const float bx = dot(d, bmX), by = dot(d, bmY);
        if (std::fabs(px) > 1.0e-3f && std::fabs(bx) > 1.0e-6f) {
            const float c = std::fabs(px * a / bx);
            if (c > 1.0f && std::isfinite(c)) ax.push_back(c);
        }
        if (std::fabs(py) > 1.0e-3f && std::fabs(by) > 1.0e-6f) {
            const float c = std::fabs(py * a / by);
            if (c > 1.0f && std::isfinite(c)) ay.push_back(c);
        }
    }
    std::sort(ax.begin(), ax.end());
    std::sort(ay.begin(), ay.end());
    sx = ax.empty() ? 0.0f : ax[ax.size() / 2];
    sy = ay.empty() ? 0.0f : ay[ay.size() / 2];
    return sx > 0.0f || sy > 0.0f;
}

static float sigma_to_mm(float lenToMm, float s) { return (lenToMm > 1.0f && s > 0.0f && s < 1.0f) ? s * 10.0f : s; }

// The CarbonPBS profile model of a beam and the hooks that apply (wrapper:2816-2862 table shape,
// 4577-4603 per-layer row and beam parameters, 4656-4660 / 4814-4819 the hooks' conditions).
static void plan_profile_hooks(const pbs_beam_settings &b, const pbs_energy_struct &en, BeamPlan &p) {
    p.profileHooks = 0;
    p.profileLayer.clear();
    const int want = b.carbonProfileHooks & 3;
    const int L = b.numLayers;
    const size_t size = b.profileData ? b.profileDataCount : 0;
    const bool model = b.profileSetting && b.numProfileSetting >= 3 && std::lround(b.profileSetting[2]) > 0 && size > 0;
    if (!want || !model) return;
    const int depthN = std::max(1, int(std::lround(b.profileSetting[2])));
    const size_t nPara = b.beamParaData ? size_t(std::max(b.numBeamParaData, 0)) : 0;
    const size_t nEn = b.profileEnergies ? size_t(std::max(b.numProfileEnergies, 0)) : 0;
    const int paraRows = int(nPara / 3);
    const size_t plane = size_t(depthN);
    int rows = 0;
    if (paraRows > 0 && nPara == size_t(paraRows) * 3 && size % (size_t(paraRows) * plane) == 0)
        rows = paraRows;
    else if (nEn > 0 && size % (nEn * plane) == 0)
        rows = int(nEn);
    else if (L > 0 && size % (size_t(L) * plane) == 0)
        rows = L;
    else if (size % plane == 0 && (size / plane) % 11 == 0)
        rows = int(size / plane / 11);  // [row][depth][11 channels] exports (wrapper:2840-2849)
    if (rows <= 0) return;
    const size_t perRow = size / size_t(rows);
    if (perRow < size_t(depthN)) return;
    const int channels = int(perRow / size_t(depthN));
    if (channels <= 0 || size_t(rows) * depthN * channels != size) return;
    const bool haveBeamPara = nPara >= size_t(L) * 3;
    int hooks = 0;
    if ((want & 1) && haveBeamPara && channels >= 3) hooks |= 1;
    if (want & 2) hooks |= 2;
    if (!hooks) return;
    p.profileHooks = hooks;
    p.profileRows = rows;
    p.profileDepthN = depthN;
    p.profileChannels = channels;
    p.profileDepth0 = b.profileSetting[0];
    p.profileDepthStep = b.profileSetting[1];
    p.beamParaPos = b.beamParaPos;
    p.profileLayer.assign(size_t(L) * 4, 0.0f);
    const float rowMax = float(std::max(0, rows - 1));
    for (int l = 0; l < L; ++l) {
        float row;
        if (rows == L)
            row = float(l);
        else if (nEn > 0 && rows == int(nEn))
            row = fractional_index_monotonic(b.profileEnergies, int(nEn), b.energies[l]);
        else if (rows == en.nEnergies && en.nEnergies > 0)
            row = std::min(std::max(p.layers[l].energyIdx, 0.0f), float(rows - 1));
        else
            row = std::min(std::max(float(l), 0.0f), rowMax);
        row = std::min(std::max(row, 0.0f), rowMax);
        float *o = &p.profileLayer[size_t(l) * 4];
        o[0] = row;
        for (int c = 0; c < 3; ++c) {
            if (nPara < size_t(rows) * 3) { o[1 + c] = 0.0f; continue; }
            const float cl = std::min(std::max(row, 0.0f), float(rows - 1));
            const int i0 = std::max(0, std::min(rows - 1, int(std::floor(cl))));
            const int i1 = std::max(0, std::min(rows - 1, i0 + 1));
            const float t = cl - float(i0);
            const float v0 = b.beamParaData[size_t(i0) * 3 + c], v1 = b.beamParaData[size_t(i1) * 3 + c];
            o[1 + c] = v0 + (v1 - v0) * t;
        }
    }
}

bool plan_beam(const int ctDims[3], const float ctRes[3], const float ctCorner[3], const int doseDims[3],
               const float doseRes[3], const float doseCorner[3], const pbs_beam_settings &b,
               const pbs_energy_struct &en, BeamPlan &p, const CpbOverride *ov, const PlanEnv *envIn,
               bool deferPlacement) {
    p = BeamPlan{};
    const PlanEnv env = envIn ? *envIn : PlanEnv{};
    if (b.numLayers <= 0 || b.numLayers > kMaxLayers) { p.error = "numLayers out of range"; return false; }
    p.numLayers = b.numLayers;
    p.subspotConv = b.useSubspotConvolution != 0;
    // units (wrapper:2728-2802)
    p.lenToMm = (ctRes[2] > 0.0f && ctRes[2] < 0.3f) ? 10.0f : 1.0f;
    p.energyDepthToMm = p.lenToMm;
    float peakMax = 0.0f;
    for (int i = 0; i < en.nEnergies; ++i) peakMax = std::max(peakMax, en.peakDepths[i]);
    if (env.energyDepthMode == 1) p.energyDepthToMm = env.energyDepthToMm;
    else if (env.energyDepthMode == 0 && p.lenToMm == 10.0f && peakMax > 200.0f) p.energyDepthToMm = 1.0f;
    const float lenToMm = p.lenToMm;

    // beam basis (wrapper:2893-2936)
    const F3 beamDir = normalize_safe(f3(b.beamDirection[0], b.beamDirection[1], b.beamDirection[2]));
    const F3 bmXW = normalize_safe(f3(b.bmXDirection[0], b.bmXDirection[1], b.bmXDirection[2]));
    const F3 bmYW = normalize_safe(f3(b.bmYDirection[0], b.bmYDirection[1], b.bmYDirection[2]));
    const F3 srcW = f3(b.sourcePosition[0], b.sourcePosition[1], b.sourcePosition[2]);
    const float sad = b.sad > 0.0f ? b.sad : b.sourceDist[0];
    p.sadInput = sad;
    const F3 basisCross = normalize_safe(cross(bmXW, bmYW));
    const bool haveCross = dot(basisCross, basisCross) > 1.0e-8f;
    const bool haveDir = dot(beamDir, beamDir) > 1.0e-8f;
    F3 bmZ = haveCross ? basisCross : beamDir;
    if (haveCross && haveDir && dot(bmZ, beamDir) < 0.0f) bmZ = scale(bmZ, -1.0f);
    if (!(dot(bmZ, bmZ) > 1.0e-8f)) bmZ = f3(0, 0, 1);
    F3 bmX = normalize_safe(sub(bmXW, scale(bmZ, dot(bmXW, bmZ))));
    if (!(dot(bmX, bmX) > 1.0e-8f)) bmX = normalize_safe(cross(bmYW, bmZ));
    if (!(dot(bmX, bmX) > 1.0e-8f)) bmX = f3(1, 0, 0);
    F3 bmY = normalize_safe(cross(bmZ, bmX));
    if (dot(bmY, bmYW) < 0.0f) bmY = scale(bmY, -1.0f);
    bmX = normalize_safe(cross(bmY, bmZ));
    const F3 gX = bmX, gY = scale(bmY, -1.0f), gZ = scale(bmZ, -1.0f);
    const F3 iso = add(scale(srcW, lenToMm), scale(bmZ, sad * lenToMm));
    p.bmX = bmX; p.bmY = bmY; p.bmZ = bmZ; p.iso = iso;
    p.g2w.m = M3{f3(gX.x, gY.x, gZ.x), f3(gX.y, gY.y, gZ.y), f3(gX.z, gY.z, gZ.z)};
    p.g2w.v = iso;
    const Aff w2g = inverse(p.g2w);

    // spot lattice and CPB / ray grid (wrapper:2951-3212)
    const bool lattice = ov ? false : build_lattice(b, p, deferPlacement && !b.spotPositionsAreIndices);
    const F3 doseRes_mm = f3(doseRes[0] * lenToMm, doseRes[1] * lenToMm, doseRes[2] * lenToMm);
    const F3 doseCorner_mm = f3(doseCorner[0] * lenToMm, doseCorner[1] * lenToMm, doseCorner[2] * lenToMm);
    if (lattice && b.raySpacing[0] > 0.0f && b.raySpacing[1] > 0.0f) {
        float msx = 0.f, msy = 0.f;
        for (int l = 0; l < b.numLayers; ++l) {
            msx = std::max(msx, b.spotSigmas[2 * l]);
            msy = std::max(msy, b.spotSigmas[2 * l + 1]);
        }
        const float mX = 3.0f * msx + 0.5f * b.raySpacing[0], mY = 3.0f * msy + 0.5f * b.raySpacing[1];
        float ox, oy;
        int nx, ny;
        if (!align_grid(p.spotOffset[0] - mX, p.spotOffset[0] + float(p.spotDims[0] - 1) * p.spotDelta[0] + mX,
                        p.spotOffset[0], b.raySpacing[0], kTileX, ox, nx) ||
            !align_grid(p.spotOffset[1] - mY, p.spotOffset[1] + float(p.spotDims[1] - 1) * p.spotDelta[1] + mY,
                        p.spotOffset[1], b.raySpacing[1], kTileY, oy, ny)) {
            p.error = "Failed to align CPB grid to raw spot lattice phase";
            return false;
        }
        p.cpbCorner[0] = ox; p.cpbCorner[1] = oy; p.cpbCorner[2] = 0.0f;
        p.cpbRes[0] = b.raySpacing[0]; p.cpbRes[1] = b.raySpacing[1]; p.cpbRes[2] = doseRes[2];
        p.rayDims[0] = nx; p.rayDims[1] = ny;
        p.legacyPath = false;
    } else if (ov) {
        p.denseWeights.clear();
        p.cpbCorner[0] = ov->corner[0]; p.cpbCorner[1] = ov->corner[1]; p.cpbCorner[2] = 0.0f;
        p.cpbRes[0] = ov->res[0]; p.cpbRes[1] = ov->res[1]; p.cpbRes[2] = doseRes[2];
        p.rayDims[0] = std::max(1, ov->dims[0]);
        p.rayDims[1] = std::max(1, ov->dims[1]);
        p.legacyPath = true;
    } else {
        if (!(b.raySpacing[0] > 0.0f) || !(b.raySpacing[1] > 0.0f)) {
            p.error = "Legacy CPB path requires explicit beam.raySpacing";
            return false;
        }
        p.denseWeights.clear();
        float minGX = 1e30f, maxGX = -1e30f, minGY = 1e30f, maxGY = -1e30f;
        const F3 sz = f3(float(std::max(doseDims[0] - 1, 0)) * doseRes_mm.x, float(std::max(doseDims[1] - 1, 0)) * doseRes_mm.y,
                         float(std::max(doseDims[2] - 1, 0)) * doseRes_mm.z);
        for (int cx = 0; cx <= 1; ++cx)
            for (int cy = 0; cy <= 1; ++cy)
                for (int cz = 0; cz <= 1; ++cz) {
                    const F3 g = apply(w2g, add(doseCorner_mm, f3(float(cx) * sz.x, float(cy) * sz.y, float(cz) * sz.z)));
                    minGX = std::fmin(minGX, g.x); maxGX = std::fmax(maxGX, g.x);
                    minGY = std::fmin(minGY, g.y); maxGY = std::fmax(maxGY, g.y);
                }
        for (int l = 0; l < b.numLayers; ++l)
            for (int s = 0; s < b.maxSubspotsPerLayer; ++s) {
                const float *q = b.subspotData + (size_t(l) * b.maxSubspotsPerLayer + s) * 5;
                if (q[2] > 0.0f) {
                    const float sx = q[0] / lenToMm, sy = q[1] / lenToMm;
                    minGX = std::fmin(minGX, sx); maxGX = std::fmax(maxGX, sx);
                    minGY = std::fmin(minGY, sy); maxGY = std::fmax(maxGY, sy);
                }
            }
        float covX = (maxGX - minGX) / lenToMm, covY = (maxGY - minGY) / lenToMm;
        if (!(covX > 0.0f) || !(covY > 0.0f)) { p.error = "Legacy CPB path produced non-positive coverage"; return false; }
        if (env.cpbCoverage > 0.0f) covX = covY = env.cpbCoverage;
        p.cpbCorner[0] = minGX / lenToMm; p.cpbCorner[1] = minGY / lenToMm; p.cpbCorner[2] = 0.0f;
        p.cpbRes[0] = b.raySpacing[0]; p.cpbRes[1] = b.raySpacing[1]; p.cpbRes[2] = doseRes[2];
        p.rayDims[0] = std::max(1, int(std::ceil(covX / p.cpbRes[0])));
        p.rayDims[1] = std::max(1, int(std::ceil(covY / p.cpbRes[1])));
        p.legacyPath = true;
    }
    p.superpDims[0] = ((p.rayDims[0] + kTileX - 1) / kTileX) * kTileX;
    p.superpDims[1] = ((p.rayDims[1] + kTileY - 1) / kTileY) * kTileY;

    // index transforms (wrapper:3233-3248)
    const F3 ctRes_mm = f3(ctRes[0] * lenToMm, ctRes[1] * lenToMm, ctRes[2] * lenToMm);
    const F3 ctCorner_mm = f3(ctCorner[0] * lenToMm, ctCorner[1] * lenToMm, ctCorner[2] * lenToMm);
    Aff w2im, w2dose;
    w2im.m = diag(1.0f / ctRes_mm.x, 1.0f / ctRes_mm.y, 1.0f / ctRes_mm.z);
    w2im.v = f3(-ctCorner_mm.x / ctRes_mm.x, -ctCorner_mm.y / ctRes_mm.y, -ctCorner_mm.z / ctRes_mm.z);
    w2dose.m = diag(1.0f / doseRes_mm.x, 1.0f / doseRes_mm.y, 1.0f / doseRes_mm.z);
    w2dose.v = f3(-doseCorner_mm.x / doseRes_mm.x, -doseCorner_mm.y / doseRes_mm.y, -doseCorner_mm.z / doseRes_mm.z);
    p.g2im = concat(p.g2w, w2im);
    p.g2dose = concat(p.g2w, w2dose);
    p.dose2g = inverse(p.g2dose);

    // source distances (wrapper:3250-3271)
    float isx = 0, isy = 0;
    const bool inferred = infer_source_dist(b, bmX, bmY, bmZ, sad, isx, isy);
    float dxmm = b.sourceDist[0] * lenToMm, dymm = b.sourceDist[1] * lenToMm;
    const bool legacySad = std::fabs(b.sourceDist[0] - b.sad) < 1.0e-3f && std::fabs(b.sourceDist[1] - b.sad) < 1.0e-3f;
    if (inferred && (!(dxmm > 1e-6f) || legacySad)) dxmm = isx * lenToMm;
    if (inferred && (!(dymm > 1e-6f) || legacySad)) dymm = isy * lenToMm;
    if (!(dxmm > 1e-6f)) dxmm = sad * lenToMm;
    if (!(dymm > 1e-6f)) dymm = sad * lenToMm;
    p.sourceDist[0] = dxmm; p.sourceDist[1] = dymm;

    // tracer extent (wrapper:3287-3370)
    float minGZ = 1e30f, maxGZ = -1e30f;
    const F3 csz = f3(float(std::max(ctDims[0] - 1, 0)) * ctRes_mm.x, float(std::max(ctDims[1] - 1, 0)) * ctRes_mm.y,
                      float(std::max(ctDims[2] - 1, 0)) * ctRes_mm.z);
    for (int cx = 0; cx <= 1; ++cx)
        for (int cy = 0; cy <= 1; ++cy)
            for (int cz = 0; cz <= 1; ++cz) {
                const F3 g = apply(w2g, add(ctCorner_mm, f3(float(cx) * csz.x, float(cy) * csz.y, float(cz) * csz.z)));
                minGZ = std::fmin(minGZ, g.z); maxGZ = std::fmax(maxGZ, g.z);
            }
    // a WEQ volume sets the step length and count and aligns the start plane to its depth_start
    // when that lands within the volume's depth span (or 50 mm) of the CT's (wrapper:3305-3370)
    p.weqVolume = b.waterEquivalence && b.waterEquivalenceCount > 9;
    float stepLen = ctRes[2] * lenToMm;
    float startZ = maxGZ;
    if (p.weqVolume) {
        const float *h = b.waterEquivalence;
        const float sl = h[1] * lenToMm;
        if (sl > 0.0f) stepLen = sl;
        const float proj = h[0] * lenToMm;
        const float aligned = 0.5f * (p.sourceDist[0] + p.sourceDist[1]) - proj;
        const float span = std::fabs(h[1] * lenToMm) * float(std::max(0, int(std::lround(h[2]))));
        if (std::isfinite(aligned) && std::isfinite(proj) && proj > 0.0f &&
            std::fabs(aligned - maxGZ) <= std::max(span, 50.0f))
            startZ = aligned;
        p.steps = std::max(1, int(std::lround(h[2])));
        const int nx = std::max(1, int(std::lround(h[8]))), ny = std::max(1, int(std::lround(h[5])));
        if (b.waterEquivalenceCount - 9 < size_t(nx) * ny * size_t(p.steps)) {
            p.error = "water_equivalence payload is smaller than header-declared nx*ny*nStep";
            return false;
        }
        p.weqDims[0] = nx; p.weqDims[1] = ny; p.weqDims[2] = p.steps;
        p.weqStepMm = std::fabs(h[1] * lenToMm);
        p.weqX0 = h[6] / lenToMm; p.weqY0 = h[3] / lenToMm;
        p.weqDx = h[7] / lenToMm; p.weqDy = h[4] / lenToMm;
    } else {
        p.steps = std::max(1, int(std::ceil((maxGZ - minGZ) / std::fabs(stepLen))) + 1);
    }
    p.spotIndices = b.spotPositionsAreIndices != 0;
    p.fanCorner[0] = p.cpbCorner[0] * lenToMm;
    p.fanCorner[1] = p.cpbCorner[1] * lenToMm;
    p.fanCorner[2] = startZ;
    p.fanDelta[0] = p.cpbRes[0] * lenToMm;
    p.fanDelta[1] = p.cpbRes[1] * lenToMm;
    p.fanDelta[2] = -std::fabs(stepLen);

    {  // frustum volume per step (wrapper:4276-4289)
        const float dX = p.fanDelta[0], dY = p.fanDelta[1], dZ = p.fanDelta[2], cz = p.fanCorner[2];
        const float sx = p.sourceDist[0], sy = p.sourceDist[1];
        const float dxyz = std::fabs(dX * dY * dZ);
        p.volConst = dxyz * (1.0f - cz / sx - cz / sy + (cz * cz + dZ * dZ / 12.0f) / (sx * sy));
        p.volLin = dxyz * dZ * (-1.0f / sx - 1.0f / sy + 2.0f * cz / (sx * sy));
        p.volSq = dxyz * dZ * dZ / (sx * sy);
    }

    // per-layer LUT parameters (conv stage wrapper:3746-3752; fill stage 3926-4329)
    float maxE = b.energies[0];
    for (int l = 1; l < b.numLayers; ++l) maxE = std::max(maxE, b.energies[l]);
    {
        const float mi = find_decimal_ordered(en.energiesPerU, en.nEnergies, maxE);
        const float maxPeak = interp_at(en.peakDepths, en.nEnergies, mi) * p.energyDepthToMm;
        p.maxCutoffWepl = 1.05f * maxPeak;
    }
    p.layers.resize(b.numLayers);
    for (int l = 0; l < b.numLayers; ++l) {
        LayerParams &L = p.layers[l];
        std::memset(&L, 0, sizeof(L));
        const float ci = find_decimal_ordered(en.energiesPerU, en.nEnergies, b.energies[l]);
        const float convPeak = interp_at(en.peakDepths, en.nEnergies, ci) * p.energyDepthToMm;
        L.convA = 0.00270f / (convPeak - 4.50f);
        L.convB = -4.39f / (convPeak - 3.86f);
        L.sigmaX = sigma_to_mm(lenToMm, b.spotSigmas[2 * l]);
        L.sigmaY = sigma_to_mm(lenToMm, b.spotSigmas[2 * l + 1]);
        const float ei = layer_energy_index(en, b.energies[l]);
        L.energyIdx = ei;
        L.scaleFact = interp_at(en.scaleFacts, en.nEnergies, ei) / p.energyDepthToMm;
        L.peakDepth = interp_at(en.peakDepths, en.nEnergies, ei) * p.energyDepthToMm;
        L.rangeStop = 1.05f * L.peakDepth;
        L.cutoffWepl = 1.05f * L.peakDepth;
        const float a = 0.00270f / (L.peakDepth - 4.50f), c = -4.39f / (L.peakDepth - 3.86f);
        L.airA = a;
        L.airC = c;
        L.airQuad = a * 1.0f * p.fanDelta[2] * p.fanDelta[2];
        L.airLin = 2.0f * a * 1.0f * p.fanDelta[2] * p.fanCorner[2] + c * p.fanDelta[2];
    }
    plan_profile_hooks(b, en, p);
    return true;
}

// One axis of the physical PB lattice of index-coded spots (buildPhysicalPBLatticeView's
// inferAxis, wrapper:1555-1634): distinct values (0.05 mm merge, first seen kept), sorted; the
// step is the median gap, every gap must be 1-32 steps within a quarter step, and every value a
// lattice point within a quarter cell.
static bool pb_axis(const std::vector<float> &vals, float &delta, float &offset, int &count) {
    if (vals.empty()) return false;
    std::vector<float> u;
    for (float v : vals) {
        bool seen = false;
        for (float w : u)
            if (std::fabs(v - w) <= 0.05f) { seen = true; break; }
        if (!seen) u.push_back(v);
    }
    std::sort(u.begin(), u.end());
    offset = u.front();
    if (u.size() == 1) {
        delta = 1.0f;
        count = 1;
        return true;
    }
    std::vector<float> gaps;
    for (size_t i = 1; i < u.size(); ++i) gaps.push_back(u[i] - u[i - 1]);
    std::vector<float> sorted = gaps;
    std::sort(sorted.begin(), sorted.end());
    const float med = sorted[sorted.size() / 2];
    if (!(med > 0.0f)) return false;
    for (float d : gaps) {
        const float ratio = d / med, k = std::round(ratio);
        if (k < 1.0f || k > 32.0f || std::fabs(ratio - k) > 0.25f) return false;
    }
    delta = med;
    count = std::max(1, int(std::lround((u.back() - u.front()) / med)) + 1);
    for (float v : vals) {
        const float f = (v - offset) / delta, k = std::round(f);
        if (k < 0.0f || k > float(count - 1) || std::fabs(f - k) > 0.25f) return false;
    }
    return true;
}

// The physical PB lattice of index-coded spots (buildPhysicalPBLatticeView, wrapper:1478-1820):
// spots decoded to mm (x = h6 + (i - 1/2) h7, y = h3 + (j - 1/2) h4), axes inferred over all
// spots, else over the layer with the most spots (offset and extent from all spots), else the WEQ
// header lattice (out-of-range spots then clamped to the edge); each weight goes to its nearest
// cell.
static bool pb_lattice_view(const pbs_beam_settings &b, HaloPlan &h, std::vector<float> &dense, std::string &error) {
    const float *hd = b.waterEquivalence;
    if (!hd || b.waterEquivalenceCount < 9) {
        error = "spotPositionsAreIndices requires a 9-value WEQ header";
        return false;
    }
    const int n = b.numSpots;
    std::vector<float> xs(n), ys(n);
    for (int i = 0; i < n; ++i) {
        xs[i] = hd[6] + (b.spotPositions[2 * i] - 0.5f) * hd[7];
        ys[i] = hd[3] + (b.spotPositions[2 * i + 1] - 0.5f) * hd[4];
    }
    int nx = 0, ny = 0;
    float dx = 0.f, dy = 0.f, ox = 0.f, oy = 0.f;
    bool fallback = false;
    if (!(pb_axis(xs, dx, ox, nx) && pb_axis(ys, dy, oy, ny))) {
        int best = -1, bestCount = 0, start = 0, bestStart = 0;
        for (int l = 0; l < b.numLayers; ++l) {
            if (b.layerSpotCounts[l] > bestCount) { bestCount = b.layerSpotCounts[l]; best = l; bestStart = start; }
            start += b.layerSpotCounts[l];
        }
        bool ok = false;
        if (best >= 0 && bestCount >= 3) {
            std::vector<float> lx(xs.begin() + bestStart, xs.begin() + bestStart + bestCount);
            std::vector<float> ly(ys.begin() + bestStart, ys.begin() + bestStart + bestCount);
            float ldx, ldy, lox, loy;
            int lnx, lny;
            if (pb_axis(lx, ldx, lox, lnx) && pb_axis(ly, ldy, loy, lny)) {
                dx = ldx;
                dy = ldy;
                const float minX = *std::min_element(xs.begin(), xs.end()), maxX = *std::max_element(xs.begin(), xs.end());
                const float minY = *std::min_element(ys.begin(), ys.end()), maxY = *std::max_element(ys.begin(), ys.end());
                nx = std::max(1, int(std::lround((maxX - minX) / dx)) + 1);
                ny = std::max(1, int(std::lround((maxY - minY) / dy)) + 1);
                ox = minX;
                oy = minY;
                ok = true;
            }
        }
        if (!ok) {
            nx = std::max(1, int(std::lround(hd[8])));
            ny = std::max(1, int(std::lround(hd[5])));
            dx = hd[7];
            dy = hd[4];
            ox = hd[6];
            oy = hd[3];
            fallback = true;
        }
    }
    if (!(dx > 0.0f) || !(dy > 0.0f) || nx <= 0 || ny <= 0) {
        error = "inferred lattice has invalid cardinality or step";
        return false;
    }
    dense.assign(size_t(nx) * ny * b.numLayers, 0.0f);
    int off = 0;
    for (int l = 0; l < b.numLayers; ++l) {
        for (int i = 0; i < b.layerSpotCounts[l]; ++i) {
            const int s = off + i;
            const float w = b.spotWeights[s];
            if (!std::isfinite(w) || w < 0.0f) {
                error = "spot weights must be finite and >= 0";
                return false;
            }
            if (w == 0.0f) continue;
            int ix = int(std::lround((xs[s] - ox) / dx)), iy = int(std::lround((ys[s] - oy) / dy));
            if (ix < 0 || ix >= nx || iy < 0 || iy >= ny) {
                if (!fallback) {
                    error = "decoded spot is outside inferred PB lattice bounds";
                    return false;
                }
                ix = std::max(0, std::min(nx - 1, ix));
                iy = std::max(0, std::min(ny - 1, iy));
            }
            dense[(size_t(l) * ny + iy) * nx + ix] += w;
        }
        off += b.layerSpotCounts[l];
    }
    h.rawDims[0] = nx; h.rawDims[1] = ny;
    h.offset[0] = ox; h.offset[1] = oy;
    h.delta[0] = dx; h.delta[1] = dy;
    h.latticeFallback = fallback;
    return true;
}

bool plan_halo(const BeamPlan &p, const pbs_beam_settings &b, HaloPlan &h, std::string &error) {
    if (p.legacyPath || p.spotDims[0] <= 0 || p.spotDims[1] <= 0) {
        error = "nuclear_correction=true requires raw spot lattice inputs; legacy CPB fallback cannot drive halo mode";
        return false;
    }
    // the halo cells are the spots' lattice: the planner's own for physical positions, the
    // physical PB view for index-coded spots (wrapper:3421-3437)
    std::vector<float> view;
    const std::vector<float> *dense = &p.denseWeights;
    if (p.spotIndices) {
        if (!pb_lattice_view(b, h, view, error)) return false;
        dense = &view;
    } else {
        h.rawDims[0] = p.spotDims[0]; h.rawDims[1] = p.spotDims[1];
        h.offset[0] = p.spotOffset[0]; h.offset[1] = p.spotOffset[1];
        h.delta[0] = p.spotDelta[0]; h.delta[1] = p.spotDelta[1];
        h.latticeFallback = false;
    }
    const int rawX = h.rawDims[0], rawY = h.rawDims[1];
    const int nucX = (rawX + 31) / 32 * 32, nucY = (rawY + 7) / 8 * 8;
    h.nucDims[0] = nucX;
    h.nucDims[1] = nucY;
    const int rx = p.rayDims[0], ry = p.rayDims[1];
    h.rayToNuc.assign(size_t(rx) * ry, -1);
    for (int j = 0; j < rawY; ++j) {
        const float gy = h.offset[1] + float(j) * h.delta[1];
        const int rayY = int(std::lround((gy - p.cpbCorner[1]) / p.cpbRes[1]));
        if (rayY < 0 || rayY >= ry) {
            error = "dense halo lattice row mapped outside active ray grid";
            return false;
        }
        for (int i = 0; i < rawX; ++i) {
            const float gx = h.offset[0] + float(i) * h.delta[0];
            const int rayX = int(std::lround((gx - p.cpbCorner[0]) / p.cpbRes[0]));
            if (rayX < 0 || rayX >= rx) {
                error = "dense halo lattice column mapped outside active ray grid";
                return false;
            }
            h.rayToNuc[size_t(rayY) * rx + rayX] = j * nucX + i;
        }
    }
    const int L = p.numLayers;
    h.weights.assign(size_t(L) * nucX * nucY, 0.0f);
    for (int l = 0; l < L; ++l)
        for (int j = 0; j < rawY; ++j)
            for (int i = 0; i < rawX; ++i)
                h.weights[(size_t(l) * nucY + j) * nucX + i] = (*dense)[(size_t(l) * rawY + j) * rawX + i];
    // spot spacing: the beam's spotDelta when set (> 1.5 rays), else the halo lattice's
    // (wrapper:4149-4171)
    const bool beamDelta = b.spotDelta[0] > b.raySpacing[0] * 1.5f;
    const float spacing = beamDelta ? b.spotDelta[0] : h.delta[0];
    if (b.raySpacing[0] > 0.0f && spacing / b.raySpacing[0] < 1.5f) {
        error = "nuclear superposition gate: spotDistInRays < 1.5 (the halo radius would exceed 32)";
        return false;
    }
    h.spotDist = std::fabs(p.fanDelta[0]) > 1e-6f ? std::fabs(spacing / p.fanDelta[0]) : 1.0f;
    return true;
}

}  // namespace pbs
