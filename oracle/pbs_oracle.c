/*
 * FP64 CPU oracle of the pencil-beam final-dose path. TEST INFRASTRUCTURE ONLY:
 * loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker; nothing that ships calls it.
 *
 * Host planning is evaluated in float in the reference's operation order (the integers it
 * yields are defined by that arithmetic); per-ray / per-step / per-pixel / per-voxel
 * dose arithmetic is evaluated in double. Citations are to files under the reference
 * repository (src/..., include/...).
 */
#ifdef _OPENMP
#include <omp.h>
#endif
#include "pbs_oracle.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define HALFF 0.5f
#define ORC_INF_F 1.0e20f /* Macro.cuh:8 INF, as float */

/* ------------------------------------------------------------------------------------
 * float geometry (include/algorithms/matrix_3x3.h, float3_affine_transform.h,
 * float3_idx_transform.h, fan_transforms.h): restated with plain float arithmetic.
 * ---------------------------------------------------------------------------------- */
typedef struct { float x, y, z; } f3;
typedef struct { f3 r0, r1, r2; } m3;
typedef struct { m3 m; f3 v; } aff;

static f3 mk3(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static f3 add3(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
static f3 sub3(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
static f3 mul3s(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
static float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static f3 cross3(f3 a, f3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* normalizeSafe lambda, wrapper:2886-2890 */
static f3 norm_safe(f3 v) {
    float len = sqrtf(dot3(v, v));
    if (len > 1e-8f) { v.x /= len; v.y /= len; v.z /= len; }
    return v;
}
static f3 mv3(m3 m, f3 a) { return mk3(dot3(m.r0, a), dot3(m.r1, a), dot3(m.r2, a)); }
static m3 mt3(m3 m) {
    m3 t;
    t.r0 = mk3(m.r0.x, m.r1.x, m.r2.x);
    t.r1 = mk3(m.r0.y, m.r1.y, m.r2.y);
    t.r2 = mk3(m.r0.z, m.r1.z, m.r2.z);
    return t;
}
/* Matrix3x3::operator*(Matrix3x3): rows of A dotted with rows of B^T (matrix_3x3.h:66-75) */
static m3 mm3(m3 a, m3 b) {
    m3 bt = mt3(b), r;
    r.r0 = mk3(dot3(a.r0, bt.r0), dot3(a.r0, bt.r1), dot3(a.r0, bt.r2));
    r.r1 = mk3(dot3(a.r1, bt.r0), dot3(a.r1, bt.r1), dot3(a.r1, bt.r2));
    r.r2 = mk3(dot3(a.r2, bt.r0), dot3(a.r2, bt.r1), dot3(a.r2, bt.r2));
    return r;
}
static m3 mdiag3(float a, float b, float c) {
    m3 r;
    r.r0 = mk3(a, 0.f, 0.f); r.r1 = mk3(0.f, b, 0.f); r.r2 = mk3(0.f, 0.f, c);
    return r;
}
/* Matrix3x3::inverse: adjugate times (1/det) as a scalar matrix (matrix_3x3.h:80-93) */
static m3 minv3(m3 m) {
    f3 r0 = m.r0, r1 = m.r1, r2 = m.r2;
    float d = (r0.x * (r1.y * r2.z - r1.z * r2.y) - r0.y * (r1.x * r2.z - r1.z * r2.x) +
               r0.z * (r1.x * r2.y - r1.y * r2.x));
    float od = 1.0f / d;
    m3 adj;
    adj.r0 = mk3(r1.y * r2.z - r1.z * r2.y, r0.z * r2.y - r0.y * r2.z, r0.y * r1.z - r0.z * r1.y);
    adj.r1 = mk3(r1.z * r2.x - r1.x * r2.z, r0.x * r2.z - r0.z * r2.x, r0.z * r1.x - r0.x * r1.z);
    adj.r2 = mk3(r1.x * r2.y - r1.y * r2.x, r0.y * r2.x - r0.x * r2.y, r0.x * r1.y - r0.y * r1.x);
    return mm3(adj, mdiag3(od, od, od));
}
static f3 aff_apply(aff t, f3 a) { return add3(mv3(t.m, a), t.v); }
static aff aff_inv(aff t) {
    aff r;
    r.m = minv3(t.m);
    r.v = mv3(r.m, mul3s(t.v, -1.0f));
    return r;
}
/* concat(first, second): apply first, then second (float3_affine_transform.h:57-61) */
static aff aff_concat(aff first, aff second) {
    aff r;
    r.m = mm3(second.m, first.m);
    r.v = add3(mv3(second.m, first.v), second.v);
    return r;
}

static void m3_to9(m3 m, float *o) {
    o[0] = m.r0.x; o[1] = m.r0.y; o[2] = m.r0.z;
    o[3] = m.r1.x; o[4] = m.r1.y; o[5] = m.r1.z;
    o[6] = m.r2.x; o[7] = m.r2.y; o[8] = m.r2.z;
}
static m3 m3_from9(const float *o) {
    m3 m;
    m.r0 = mk3(o[0], o[1], o[2]); m.r1 = mk3(o[3], o[4], o[5]); m.r2 = mk3(o[6], o[7], o[8]);
    return m;
}

/* ------------------------------------------------------------------------------------
 * ordered-vector helpers (include/utils/vector_find.h, vector_interpolate.h)
 * ---------------------------------------------------------------------------------- */
static int last_smaller_or_eq_f(const float *v, int n, float value) {
    int upper = n - 1, lower = 0;
    if (v[n - 1] <= value) return upper;
    if (v[0] > value) return -1;
    while (upper - lower > 1) {
        int pivot = (upper + lower) / 2;
        if (v[pivot] <= value) lower = pivot; else upper = pivot;
    }
    return lower;
}

float orc_find_decimal_ordered(const float *v, int n, float value) {
    if (value >= v[n - 1]) return (float)(n - 1);
    if (value < v[0]) return 0.0f;
    unsigned floorIdx = (unsigned)last_smaller_or_eq_f(v, n, value);
    float corr = (value - v[floorIdx]) / (v[floorIdx + 1] - v[floorIdx]);
    return (float)floorIdx + corr;
}

float orc_vector_interpolate(const float *v, int n, float idx) {
    if (idx <= 0.0f) return v[0];
    if (idx >= (float)(n - 1)) return v[n - 1];
    float ip;
    float dec = modff(idx, &ip);
    unsigned fi = (unsigned)ip;
    float corr = (v[fi + 1] - v[fi]) * dec;
    return v[fi] + corr;
}

int orc_find_first_larger_ordered(const double *v, int n, double value) {
    int upper = n - 1, lower = 0;
    if (v[n - 1] <= value) return upper;
    if (v[0] > value) return 0;
    while (upper - lower > 1) {
        int pivot = (upper + lower) / 2;
        if (v[pivot] <= value) lower = pivot; else upper = pivot;
    }
    return lower + 1;
}

int orc_align_grid(float minC, float maxC, float ref, float step, int multiple,
                   float *origin, int *count, int *first, int *last) {
    if (!(step > 0.0f) || !isfinite(minC) || !isfinite(maxC) || !isfinite(ref)) return 0;
    if (maxC < minC) { float t = minC; minC = maxC; maxC = t; }
    const float eps = 1.0e-4f;
    int fs = (int)ceilf((minC - ref) / step - eps);
    int ls = (int)floorf((maxC - ref) / step + eps);
    if (ls < fs) return 0;
    *origin = ref + (float)fs * step;
    int c = ls - fs + 1;
    if (c < 1) c = 1;
    if (multiple > 1) c = ((c + multiple - 1) / multiple) * multiple;
    *count = c;
    if (first) *first = fs;
    if (last) *last = ls;
    return 1;
}

int orc_tile_radius_f32(float m) {
    /* complete_superposition.cu:242: min(int(KS_SIGMA_CUTOFF/(sqrt(2)*min) + 0.5), 33).
     * A NaN minimum (no finite, no +inf entry) is mapped to radius 0. */
    if (m != m) return 0;
    float q = 3.0f / (sqrtf(2.0f) * m) + HALFF;
    int r = (int)q;
    return r < ORC_MAX_RAD + 1 ? r : ORC_MAX_RAD + 1;
}

void orc_batch_bins(const int *histIn, int *batched, int *effRad) {
    int hist[ORC_MAX_RAD + 2];
    memcpy(hist, histIn, sizeof(hist));
    /* primary overflow policy: clamp bin 33 into bin 32 (complete_superposition.cu:413-463) */
    hist[ORC_MAX_RAD] += hist[ORC_MAX_RAD + 1];
    hist[ORC_MAX_RAD + 1] = 0;
    int layerMax = 0;
    for (int i = 0; i < ORC_MAX_RAD + 2; ++i) if (hist[i] > 0) layerMax = i;
    for (int i = 0; i <= ORC_MAX_RAD; ++i) batched[i] = 0;
    batched[0] = hist[0];
    int rec = layerMax;
    for (int i = 0; i < ORC_MAX_RAD + 2; ++i) effRad[i] = i;
    effRad[ORC_MAX_RAD + 1] = ORC_MAX_RAD;
    /* complete_superposition.cu:482-490: merge bins downwards until a batch holds >= 16 tiles;
     * every tile of a merged bin is superposed with the batch's (largest) radius */
    for (int rad = layerMax; rad > 0; --rad) {
        batched[rec] += hist[rad];
        effRad[rad] = rec;
        if (batched[rec] >= ORC_MIN_TILES_IN_BATCH) rec = rad - 1;
    }
}

void orc_tile_radii_f32(const float *rs, int rx, int ry, int k0, int k1, int *rad) {
    int ntx = rx / ORC_TILE_X, nty = ry / ORC_TILE_Y;
    for (int k = k0; k < k1; ++k)
        for (int ty = 0; ty < nty; ++ty)
            for (int tx = 0; tx < ntx; ++tx) {
                float m = NAN;
                for (int y = 0; y < ORC_TILE_Y; ++y)
                    for (int x = 0; x < ORC_TILE_X; ++x) {
                        float v = rs[((size_t)k * ry + ty * ORC_TILE_Y + y) * rx + tx * ORC_TILE_X + x];
                        m = fminf(m, v);
                    }
                rad[((size_t)(k - k0) * nty + ty) * ntx + tx] = orc_tile_radius_f32(m);
            }
}

/* ------------------------------------------------------------------------------------
 * sampling helpers: texture linear filtering restated (texel centres at i+0.5 in
 * texture coordinates; clamp or border-0 addressing).
 * ---------------------------------------------------------------------------------- */
static double lut1d_clamp(const float *t, int n, double xb /* texel index space */) {
    double f = floor(xb);
    double a = xb - f;
    long i0 = (long)f, i1 = i0 + 1;
    if (i0 < 0) i0 = 0; if (i0 > n - 1) i0 = n - 1;
    if (i1 < 0) i1 = 0; if (i1 > n - 1) i1 = n - 1;
    return (1.0 - a) * t[i0] + a * t[i1];
}

static double lut2d_clamp(const float *t, int nx, int ny, double xb, double yb) {
    double fx = floor(xb), fy = floor(yb);
    double ax = xb - fx, ay = yb - fy;
    long x0 = (long)fx, x1 = x0 + 1, y0 = (long)fy, y1 = y0 + 1;
    if (x0 < 0) x0 = 0; if (x0 > nx - 1) x0 = nx - 1;
    if (x1 < 0) x1 = 0; if (x1 > nx - 1) x1 = nx - 1;
    if (y0 < 0) y0 = 0; if (y0 > ny - 1) y0 = ny - 1;
    if (y1 < 0) y1 = 0; if (y1 > ny - 1) y1 = ny - 1;
    double v0 = (1.0 - ax) * t[y0 * nx + x0] + ax * t[y0 * nx + x1];
    double v1 = (1.0 - ax) * t[y1 * nx + x0] + ax * t[y1 * nx + x1];
    return (1.0 - ay) * v0 + ay * v1;
}

double orc_sample3d_border(const float *vol, const int dims[3], double x, double y, double z) {
    double fx = floor(x), fy = floor(y), fz = floor(z);
    long x0 = (long)fx, y0 = (long)fy, z0 = (long)fz;
    double tx = x - fx, ty = y - fy, tz = z - fz;
    double c[2][2][2];
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                long xi = x0 + dx, yi = y0 + dy, zi = z0 + dz;
                c[dz][dy][dx] = (xi < 0 || yi < 0 || zi < 0 || xi >= dims[0] || yi >= dims[1] || zi >= dims[2])
                                    ? 0.0 : (double)vol[((size_t)zi * dims[1] + yi) * dims[0] + xi];
            }
    double c00 = c[0][0][0] + (c[0][0][1] - c[0][0][0]) * tx;
    double c10 = c[0][1][0] + (c[0][1][1] - c[0][1][0]) * tx;
    double c01 = c[1][0][0] + (c[1][0][1] - c[1][0][0]) * tx;
    double c11 = c[1][1][0] + (c[1][1][1] - c[1][1][0]) * tx;
    double c0 = c00 + (c10 - c00) * ty;
    double c1 = c01 + (c11 - c01) * ty;
    return c0 + (c1 - c0) * tz;
}

static double sample3d_border_d(const double *vol, const int dims[3], double x, double y, double z) {
    double fx = floor(x), fy = floor(y), fz = floor(z);
    long x0 = (long)fx, y0 = (long)fy, z0 = (long)fz;
    double tx = x - fx, ty = y - fy, tz = z - fz;
    double c[2][2][2];
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                long xi = x0 + dx, yi = y0 + dy, zi = z0 + dz;
                c[dz][dy][dx] = (xi < 0 || yi < 0 || zi < 0 || xi >= dims[0] || yi >= dims[1] || zi >= dims[2])
                                    ? 0.0 : vol[((size_t)zi * dims[1] + yi) * dims[0] + xi];
            }
    double c00 = c[0][0][0] + (c[0][0][1] - c[0][0][0]) * tx;
    double c10 = c[0][1][0] + (c[0][1][1] - c[0][1][0]) * tx;
    double c01 = c[1][0][0] + (c[1][0][1] - c[1][0][0]) * tx;
    double c11 = c[1][1][0] + (c[1][1][1] - c[1][1][0]) * tx;
    double c0 = c00 + (c10 - c00) * ty;
    double c1 = c01 + (c11 - c01) * ty;
    return c0 + (c1 - c0) * tz;
}

/* ------------------------------------------------------------------------------------
 * planning (wrapper:2446-3406)
 * ---------------------------------------------------------------------------------- */
static int fail(orc_plan *p, const char *msg) {
    p->ok = 0;
    snprintf(p->error, sizeof(p->error), "%s", msg);
    return -1;
}

static int cmp_float(const void *a, const void *b) {
    float x = *(const float *)a, y = *(const float *)b;
    return (x > y) - (x < y);
}

/* buildRawSpotLattice, physical-position mode (wrapper:1259-1473). Returns dense weights
 * [L][ny][nx] (malloc'd) or NULL when the lattice cannot be built. */
static float *build_lattice(const orc_beam *b, orc_plan *p) {
    if (!b->layerSpotCounts || !b->spotPositions || !b->spotWeights || b->numSpots <= 0) return NULL;
    int total = 0;
    for (int l = 0; l < b->numLayers; ++l) total += b->layerSpotCounts[l];
    if (total <= 0 || total != b->numSpots) return NULL;
    if (b->spotIndices) {
        /* index mode (wrapper:1295-1364): the WEQ ray lattice of the header; each spot split
         * bilinearly at its texel coordinate - 1/2 over the four cells around it (wrapper:931-968) */
        if (!b->weq || b->weqCount < 9) return NULL;
        const float *h = b->weq;
        int nx = (int)lroundf(h[8]), ny = (int)lroundf(h[5]);
        if (nx < 1) nx = 1;
        if (ny < 1) ny = 1;
        float *dense = (float *)calloc((size_t)nx * ny * b->numLayers, sizeof(float));
        int off = 0;
        for (int l = 0; l < b->numLayers; ++l) {
            int cnt = b->layerSpotCounts[l];
            if (cnt <= 0) { free(dense); return NULL; }
            for (int i = 0; i < cnt; ++i) {
                int s = off + i;
                float w = b->spotWeights[s];
                if (!isfinite(w) || w < 0.0f) { free(dense); return NULL; }
                if (w == 0.0f) continue;
                float lx = b->spotPositions[2 * s] - 0.5f, ly = b->spotPositions[2 * s + 1] - 0.5f;
                int x0 = (int)floorf(lx), y0 = (int)floorf(ly);
                float fx = lx - (float)x0, fy = ly - (float)y0;
                float frac[4] = {(1.0f - fx) * (1.0f - fy), fx * (1.0f - fy), (1.0f - fx) * fy, fx * fy};
                int any = 0;
                for (int q = 0; q < 4; ++q) {
                    int cx = x0 + (q & 1), cy = y0 + (q >> 1);
                    if (!(frac[q] > 0.0f) || cx < 0 || cx >= nx || cy < 0 || cy >= ny) continue;
                    dense[((size_t)l * ny + cy) * nx + cx] += w * frac[q];
                    any = 1;
                }
                if (!any) { free(dense); return NULL; }
            }
            off += cnt;
        }
        p->spotDims[0] = nx; p->spotDims[1] = ny;
        p->spotOffset[0] = h[6]; p->spotOffset[1] = h[3];
        p->spotDelta[0] = h[7]; p->spotDelta[1] = h[4];
        return dense;
    }
    float *xs = (float *)malloc(sizeof(float) * total), *ys = (float *)malloc(sizeof(float) * total);
    for (int i = 0; i < total; ++i) { xs[i] = b->spotPositions[2 * i]; ys[i] = b->spotPositions[2 * i + 1]; }
    qsort(xs, total, sizeof(float), cmp_float);
    qsort(ys, total, sizeof(float), cmp_float);
    int nux = 0, nuy = 0;
    for (int i = 0; i < total; ++i) if (nux == 0 || !(fabsf(xs[nux - 1] - xs[i]) <= 1e-4f)) xs[nux++] = xs[i];
    for (int i = 0; i < total; ++i) if (nuy == 0 || !(fabsf(ys[nuy - 1] - ys[i]) <= 1e-4f)) ys[nuy++] = ys[i];
    float dx = ORC_INF_F, dy = ORC_INF_F;
    for (int i = 1; i < nux; ++i) { float d = xs[i] - xs[i - 1]; if (d > 1.0e-4f && d < dx) dx = d; }
    for (int i = 1; i < nuy; ++i) { float d = ys[i] - ys[i - 1]; if (d > 1.0e-4f && d < dy) dy = d; }
    if (dx == ORC_INF_F) dx = 1.0f;
    if (dy == ORC_INF_F) dy = 1.0f;
    float ox = xs[0], oy = ys[0];
    int nx = (int)lroundf((xs[nux - 1] - ox) / dx) + 1;
    int ny = (int)lroundf((ys[nuy - 1] - oy) / dy) + 1;
    if (nx < 1) nx = 1;
    if (ny < 1) ny = 1;
    free(xs); free(ys);
    float *dense = (float *)calloc((size_t)nx * ny * b->numLayers, sizeof(float));
    int off = 0;
    for (int l = 0; l < b->numLayers; ++l) {
        int cnt = b->layerSpotCounts[l];
        if (cnt <= 0) { free(dense); return NULL; }
        for (int i = 0; i < cnt; ++i) {
            int s = off + i;
            float x = b->spotPositions[2 * s], y = b->spotPositions[2 * s + 1], w = b->spotWeights[s];
            if (!isfinite(w) || w < 0.0f) { free(dense); return NULL; }
            if (w == 0.0f) continue;
            float fx = (dx != 0.0f) ? ((x - ox) / dx) : 0.0f;
            float fy = (dy != 0.0f) ? ((y - oy) / dy) : 0.0f;
            int ix = (int)lroundf(fx), iy = (int)lroundf(fy);
            if (ix < 0 || ix >= nx || iy < 0 || iy >= ny) { free(dense); return NULL; }
            if (!(fabsf(ox + ix * dx - x) <= 2e-2f) || !(fabsf(oy + iy * dy - y) <= 2e-2f)) {
                free(dense); return NULL;
            }
            dense[((size_t)l * ny + iy) * nx + ix] += w;
        }
        off += cnt;
    }
    p->spotDims[0] = nx; p->spotDims[1] = ny;
    p->spotOffset[0] = ox; p->spotOffset[1] = oy;
    p->spotDelta[0] = dx; p->spotDelta[1] = dy;
    return dense;
}

static int cmp_float_sort(const void *a, const void *b) { return cmp_float(a, b); }

/* estimateVirtualSourceDistancesFromSpots (wrapper:1847-1921), physical positions */
static int infer_source_dist(const orc_beam *b, f3 bmX, f3 bmY, f3 bmZ, float sad, float *sx, float *sy) {
    if (!(sad > 1.0e-6f) || !b->spotBeamDirections || !b->spotPositions || b->numSpots <= 0) return 0;
    float *ax = (float *)malloc(sizeof(float) * b->numSpots), *ay = (float *)malloc(sizeof(float) * b->numSpots);
    int nx = 0, ny = 0;
    if (b->spotIndices && (!b->weq || b->weqCount < 9)) { free(ax); free(ay); return 0; }
    for (int i = 0; i < b->numSpots; ++i) {
        float px = b->spotPositions[2 * i], py = b->spotPositions[2 * i + 1];
        if (b->spotIndices) { /* texel coordinate -> reference plane (wrapper:1881-1888) */
            px = b->weq[6] + (px - 0.5f) * b->weq[7];
            py = b->weq[3] + (py - 0.5f) * b->weq[4];
        }
        f3 d = mk3(b->spotBeamDirections[3 * i], b->spotBeamDirections[3 * i + 1], b->spotBeamDirections[3 * i + 2]);
        float len = sqrtf(dot3(d, d));
        if (!(len > 1.0e-6f)) continue;
        d.x /= len; d.y /= len; d.z /= len;
        float a = dot3(d, bmZ);
        if (!(fabsf(a) > 1.0e-6f)) continue;
        float bx = dot3(d, bmX), by = dot3(d, bmY);
        if (fabsf(px) > 1.0e-3f && fabsf(bx) > 1.0e-6f) {
            float c = fabsf(px * a / bx);
            if (c > 1.0f && isfinite(c)) ax[nx++] = c;
        }
        if (fabsf(py) > 1.0e-3f && fabsf(by) > 1.0e-6f) {
            float c = fabsf(py * a / by);
            if (c > 1.0f && isfinite(c)) ay[ny++] = c;
        }
    }
    qsort(ax, nx, sizeof(float), cmp_float_sort);
    qsort(ay, ny, sizeof(float), cmp_float_sort);
    *sx = nx ? ax[nx / 2] : 0.0f;
    *sy = ny ? ay[ny / 2] : 0.0f;
    free(ax); free(ay);
    return (*sx > 0.0f || *sy > 0.0f);
}

/* sizeof of the structs as compiled (the Python mirrors are checked against it) */
static float frac_index_monotonic(const float *v, int n, float value) {
    /* findFractionalIndexMonotonic (wrapper:2337-2375) */
    if (n <= 1) return 0.0f;
    int inc = v[n - 1] >= v[0];
    if (inc ? value <= v[0] : value >= v[0]) return 0.0f;
    if (inc ? value >= v[n - 1] : value <= v[n - 1]) return (float)(n - 1);
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        int keepLo = inc ? v[mid] <= value : v[mid] >= value;
        if (keepLo) lo = mid; else hi = mid;
    }
    float den = v[lo + 1] - v[lo];
    float t = den != 0.0f ? (value - v[lo]) / den : 0.0f;
    if (t < 0.0f) t = 0.0f;
    if (t > 1.0f) t = 1.0f;
    return (float)lo + t;
}

int orc_ct_adapter(const float *in, size_t n, int forcedType, const orc_energy *en, float *out) {
    int type = forcedType;
    if (type < 0) {
        /* sampleCtStats / autoDetectCtType (wrapper:2307-2335): every (n / 1e5)-th value */
        size_t stride = n > 100000 ? n / 100000 : 1;
        float mn = in[0], mx = in[0];
        for (size_t i = 0; i < n; i += stride) {
            if (in[i] < mn) mn = in[i];
            if (in[i] > mx) mx = in[i];
        }
        type = mn < -200.0f ? 0 : (mx <= 20.0f && mn >= 0.0f) ? 2 : 1;
    }
    if (type == 1) {
        memcpy(out, in, n * sizeof(float));
        return 1;
    }
    if (type == 0) {
        for (size_t i = 0; i < n; ++i) out[i] = in[i] + 1000.0f;
        return 0;
    }
    const float *lut = type == 2 ? en->density : en->sp;
    int nl = type == 2 ? en->nDensity : en->nSp;
    float scale = type == 2 ? en->densityScale : en->spScale;
    if (!lut || nl <= 0 || scale == 0.0f) return -1;
    for (size_t i = 0; i < n; ++i) out[i] = frac_index_monotonic(lut, nl, in[i]) / scale;
    return type;
}

void orc_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

static float layer_energy_index(const orc_energy *en, float energy);

/* The profile model of a beam and the hooks that apply (wrapper:2816-2862 table shape; 4577-4603
 * per-layer row and beam parameters; 4656-4660 and 4814-4819 the hooks' conditions). */
static void plan_profile(const orc_beam *b, const orc_energy *en, orc_plan *p) {
    p->profileHooks = 0;
    const int want = b->carbonProfileHooks & 3;
    const int L = b->numLayers;
    const size_t size = b->profileData ? b->profileDataCount : 0;
    if (!want || !b->profileSetting || b->numProfileSetting < 3 || lroundf(b->profileSetting[2]) <= 0 || size == 0)
        return;
    const int depthN = lroundf(b->profileSetting[2]) > 1 ? (int)lroundf(b->profileSetting[2]) : 1;
    const size_t nPara = b->beamParaData ? (size_t)(b->numBeamParaData > 0 ? b->numBeamParaData : 0) : 0;
    const size_t nEn = b->profileEnergies ? (size_t)(b->numProfileEnergies > 0 ? b->numProfileEnergies : 0) : 0;
    const int paraRows = (int)(nPara / 3);
    int rows = 0;
    if (paraRows > 0 && nPara == (size_t)paraRows * 3 && size % ((size_t)paraRows * depthN) == 0) rows = paraRows;
    else if (nEn > 0 && size % (nEn * depthN) == 0) rows = (int)nEn;
    else if (L > 0 && size % ((size_t)L * depthN) == 0) rows = L;
    else if (size % depthN == 0 && (size / depthN) % 11 == 0) rows = (int)(size / depthN / 11);
    if (rows <= 0) return;
    const size_t perRow = size / (size_t)rows;
    if (perRow < (size_t)depthN) return;
    const int channels = (int)(perRow / (size_t)depthN);
    if (channels <= 0 || (size_t)rows * depthN * channels != size) return;
    int hooks = 0;
    if ((want & 1) && nPara >= (size_t)L * 3 && channels >= 3) hooks |= 1;
    if (want & 2) hooks |= 2;
    if (!hooks) return;
    p->profileHooks = hooks;
    p->profileRows = rows;
    p->profileDepthN = depthN;
    p->profileChannels = channels;
    p->profileDepth0 = b->profileSetting[0];
    p->profileDepthStep = b->profileSetting[1];
    p->beamParaPos = b->beamParaPos;
    p->profileTable = b->profileData;
    const float rowMax = (float)(rows > 1 ? rows - 1 : 0);
    for (int l = 0; l < L && l < ORC_MAX_LAYERS; ++l) {
        float row;
        if (rows == L) row = (float)l;
        else if (nEn > 0 && rows == (int)nEn) row = frac_index_monotonic(b->profileEnergies, (int)nEn, b->energies[l]);
        else if (rows == en->nEnergies && en->nEnergies > 0) {
            row = layer_energy_index(en, b->energies[l]);
            row = fminf(fmaxf(row, 0.0f), (float)(rows - 1));
        } else row = fminf(fmaxf((float)l, 0.0f), rowMax);
        row = fminf(fmaxf(row, 0.0f), rowMax);
        p->profileLayer[l][0] = row;
        for (int c = 0; c < 3; ++c) {
            if (nPara < (size_t)rows * 3) { p->profileLayer[l][1 + c] = 0.0f; continue; }
            const float cl = fminf(fmaxf(row, 0.0f), (float)(rows - 1));
            int i0 = (int)floorf(cl);
            i0 = i0 < 0 ? 0 : (i0 > rows - 1 ? rows - 1 : i0);
            const int i1 = i0 + 1 > rows - 1 ? rows - 1 : i0 + 1;
            const float t = cl - (float)i0;
            const float v0 = b->beamParaData[(size_t)i0 * 3 + c], v1 = b->beamParaData[(size_t)i1 * 3 + c];
            p->profileLayer[l][1 + c] = v0 + (v1 - v0) * t;
        }
    }
}

static double profile_at(const orc_plan *p, int c, double d, double r) {
    const int d0 = (int)floor(d), r0 = (int)floor(r);
    const int d1 = d0 + 1 < p->profileDepthN ? d0 + 1 : p->profileDepthN - 1;
    const int r1 = r0 + 1 < p->profileRows ? r0 + 1 : p->profileRows - 1;
    const double fd = d - d0, fr = r - r0;
    const int ch = p->profileChannels, dn = p->profileDepthN;
    const float *t = p->profileTable;
#define PT(rr, dd) ((double)t[((size_t)(rr) * dn + (dd)) * ch + c])
    const double v0 = PT(r0, d0) + (PT(r0, d1) - PT(r0, d0)) * fd;
    const double v1 = PT(r1, d0) + (PT(r1, d1) - PT(r1, d0)) * fd;
#undef PT
    return v0 + (v1 - v0) * fr;
}

void orc_profile_hooks(const orc_plan *p, int l, const double *wepl, double *idd, double *rsigma) {
    if (!p->profileHooks) return;
    const int rx = p->rayDims[0], ry = p->rayDims[1];
    const size_t plane = (size_t)rx * ry;
    const int k0 = p->beamFirstInside > 0 ? p->beamFirstInside : 0;
    const int k1 = p->afterLast[l] < p->steps ? p->afterLast[l] : p->steps;
    const int nG = (p->profileChannels - 1) / 2;
    const double row = p->profileLayer[l][0];
    const double sad = 0.5 * ((double)p->sourceDist[0] + (double)p->sourceDist[1]);
    const double dx = p->sourceDist[0], dy = p->sourceDist[1];
#pragma omp parallel for schedule(static)
    for (int y = 0; y < ry; ++y)
        for (int x = 0; x < rx; ++x) {
            const double fanX = (double)p->fanCorner[0] + x * (double)p->fanDelta[0];
            const double fanY = (double)p->fanCorner[1] + y * (double)p->fanDelta[1];
            const double slant = sqrt(1.0 + fanX * fanX / (dx * dx) + fanY * fanY / (dy * dy));
            for (int k = k0; k < k1; ++k) {
                const size_t i = (size_t)k * plane + (size_t)y * rx + x;
                if (!(idd[i] > 0.0)) {
                    if (p->profileHooks & 1) rsigma[i] = INFINITY;
                    continue;
                }
                double d = (wepl[i] - p->profileDepth0) / p->profileDepthStep;
                d = d < 0.0 ? 0.0 : (d > p->profileDepthN - 1 ? p->profileDepthN - 1 : d);
                if (p->profileHooks & 1) {
                    double sw = 0.0, s2 = 0.0;
                    for (int q = 0; q < nG; ++q) {
                        const double wq = profile_at(p, q, d, row), sq = profile_at(p, q + nG, d, row);
                        sw += wq;
                        s2 += wq * sq * sq;
                    }
                    if (!(sw > 1.0e-8)) {
                        rsigma[i] = INFINITY;
                    } else {
                        /* geometric depth along the divergent ray from the virtual source */
                        const double fanZ = (double)p->fanCorner[2] + k * (double)p->fanDelta[2];
                        const double phy = (sad - fanZ) * slant - (sad - p->beamParaPos);
                        const double r2 = p->profileLayer[l][1] + 2.0 * p->profileLayer[l][2] * phy +
                                          p->profileLayer[l][3] * phy * phy;
                        const double a2 = 0.5 * fmax(0.0, r2 + s2 / sw);
                        const double axis = sqrt(a2);
                        const double width = 0.5 * ((double)p->fanDelta[0] * (1.0 - fanZ / dx) +
                                                    (double)p->fanDelta[1] * (1.0 - fanZ / dy));
                        rsigma[i] = 0.5 * width / (sqrt(2.0) * (axis + 0.21));
                    }
                }
                if (p->profileHooks & 2) {
                    const double w = profile_at(p, p->profileChannels - 1, d, row);
                    idd[i] *= w > 0.0 ? w : 0.0;
                }
            }
        }
}

long orc_sizeof(int which) {
    return which == 0 ? (long)sizeof(orc_beam) : which == 1 ? (long)sizeof(orc_plan)
                      : which == 2 ? (long)sizeof(orc_energy) : -1;
}

int orc_plan_beam(const int ctDims[3], const float ctRes[3], const float ctCorner[3],
                  const int doseDims[3], const float doseRes[3], const float doseCorner[3],
                  const orc_beam *b, const orc_energy *en, orc_plan *p) {
    memset(p, 0, sizeof(*p));
    p->numLayers = b->numLayers;
    if (b->numLayers <= 0 || b->numLayers > ORC_MAX_LAYERS) return fail(p, "numLayers out of range");
    /* units (wrapper:2728-2802): lenToMm from ctResolution.z; depth tables in mm unless the
     * CT is cm-scaled and peak depths are clearly mm. */
    p->lenToMm = (ctRes[2] > 0.0f && ctRes[2] < 0.3f) ? 10.0f : 1.0f;
    p->energyDepthToMm = p->lenToMm;
    float peakMax = 0.0f;
    for (int i = 0; i < en->nEnergies; ++i) if (en->peaks[i] > peakMax) peakMax = en->peaks[i];
    if (b->energyDepthMode == 1) p->energyDepthToMm = b->energyDepthToMm;
    else if (b->energyDepthMode == 0 && p->lenToMm == 10.0f && peakMax > 200.0f) p->energyDepthToMm = 1.0f;
    const float lenToMm = p->lenToMm;

    /* beam basis (wrapper:2893-2936) */
    f3 beamDir = norm_safe(mk3(b->beamDirection[0], b->beamDirection[1], b->beamDirection[2]));
    f3 bmXW = norm_safe(mk3(b->bmXDirection[0], b->bmXDirection[1], b->bmXDirection[2]));
    f3 bmYW = norm_safe(mk3(b->bmYDirection[0], b->bmYDirection[1], b->bmYDirection[2]));
    f3 srcW = mk3(b->sourcePosition[0], b->sourcePosition[1], b->sourcePosition[2]);
    float sad = (b->sad > 0.0f) ? b->sad : b->sourceDist[0];
    f3 basisCross = norm_safe(cross3(bmXW, bmYW));
    int haveCross = dot3(basisCross, basisCross) > 1.0e-8f;
    int haveDir = dot3(beamDir, beamDir) > 1.0e-8f;
    f3 bmZ = haveCross ? basisCross : beamDir;
    if (haveCross && haveDir && dot3(bmZ, beamDir) < 0.0f) bmZ = mul3s(bmZ, -1.0f);
    if (!(dot3(bmZ, bmZ) > 1.0e-8f)) bmZ = mk3(0.f, 0.f, 1.f);
    f3 bmX = norm_safe(sub3(bmXW, mul3s(bmZ, dot3(bmXW, bmZ))));
    if (!(dot3(bmX, bmX) > 1.0e-8f)) bmX = norm_safe(cross3(bmYW, bmZ));
    if (!(dot3(bmX, bmX) > 1.0e-8f)) bmX = mk3(1.f, 0.f, 0.f);
    f3 bmY = norm_safe(cross3(bmZ, bmX));
    if (dot3(bmY, bmYW) < 0.0f) bmY = mul3s(bmY, -1.0f);
    bmX = norm_safe(cross3(bmY, bmZ));
    f3 gX = bmX, gY = mul3s(bmY, -1.0f), gZ = mul3s(bmZ, -1.0f);
    f3 src_mm = mul3s(srcW, lenToMm);
    float sad_mm = sad * lenToMm;
    f3 iso = add3(src_mm, mul3s(bmZ, sad_mm));
    p->bmX[0] = bmX.x; p->bmX[1] = bmX.y; p->bmX[2] = bmX.z;
    p->bmY[0] = bmY.x; p->bmY[1] = bmY.y; p->bmY[2] = bmY.z;
    p->bmZ[0] = bmZ.x; p->bmZ[1] = bmZ.y; p->bmZ[2] = bmZ.z;
    p->iso[0] = iso.x; p->iso[1] = iso.y; p->iso[2] = iso.z;

    aff g2w;
    g2w.m.r0 = mk3(gX.x, gY.x, gZ.x);
    g2w.m.r1 = mk3(gX.y, gY.y, gZ.y);
    g2w.m.r2 = mk3(gX.z, gY.z, gZ.z);
    g2w.v = iso;
    aff w2g = aff_inv(g2w);
    m3_to9(g2w.m, p->g2w);
    p->g2wOff[0] = iso.x; p->g2wOff[1] = iso.y; p->g2wOff[2] = iso.z;

    /* spot lattice and CPB / ray grid (wrapper:2951-3212) */
    float *lattice = build_lattice(b, p);
    const int haveLattice = lattice != NULL && !b->cpbOverride;
    free(lattice);
    f3 doseRes_mm = mk3(doseRes[0] * lenToMm, doseRes[1] * lenToMm, doseRes[2] * lenToMm);
    f3 doseCorner_mm = mk3(doseCorner[0] * lenToMm, doseCorner[1] * lenToMm, doseCorner[2] * lenToMm);
    if (haveLattice && b->raySpacing[0] > 0.0f && b->raySpacing[1] > 0.0f) {
        float msx = 0.f, msy = 0.f;
        for (int l = 0; l < b->numLayers; ++l) {
            if (b->spotSigmas[2 * l] > msx) msx = b->spotSigmas[2 * l];
            if (b->spotSigmas[2 * l + 1] > msy) msy = b->spotSigmas[2 * l + 1];
        }
        float mX = 3.0f * msx + HALFF * b->raySpacing[0];
        float mY = 3.0f * msy + HALFF * b->raySpacing[1];
        float ox, oy; int nx, ny;
        if (!orc_align_grid(p->spotOffset[0] - mX,
                            p->spotOffset[0] + (float)(p->spotDims[0] - 1) * p->spotDelta[0] + mX,
                            p->spotOffset[0], b->raySpacing[0], ORC_TILE_X, &ox, &nx, NULL, NULL) ||
            !orc_align_grid(p->spotOffset[1] - mY,
                            p->spotOffset[1] + (float)(p->spotDims[1] - 1) * p->spotDelta[1] + mY,
                            p->spotOffset[1], b->raySpacing[1], ORC_TILE_Y, &oy, &ny, NULL, NULL))
            return fail(p, "Failed to align CPB grid to raw spot lattice phase");
        p->cpbCorner[0] = ox; p->cpbCorner[1] = oy; p->cpbCorner[2] = 0.0f;
        p->cpbRes[0] = b->raySpacing[0]; p->cpbRes[1] = b->raySpacing[1]; p->cpbRes[2] = doseRes[2];
        p->rayDims[0] = nx; p->rayDims[1] = ny;
        p->legacyPath = 0;
    } else if (b->cpbOverride) {
        /* cudaFinalDose: the CPB grid is the ROI box (cuda_final_dose.cu:302-308) */
        p->cpbCorner[0] = b->cpbOvCorner[0]; p->cpbCorner[1] = b->cpbOvCorner[1]; p->cpbCorner[2] = 0.0f;
        p->cpbRes[0] = b->cpbOvRes[0]; p->cpbRes[1] = b->cpbOvRes[1]; p->cpbRes[2] = doseRes[2];
        p->rayDims[0] = b->cpbOvDims[0] > 1 ? b->cpbOvDims[0] : 1;
        p->rayDims[1] = b->cpbOvDims[1] > 1 ? b->cpbOvDims[1] : 1;
        p->legacyPath = 1;
    } else {
        /* legacy CPB path (wrapper:3175-3212): grid covers the projected dose box and subspots */
        if (!(b->raySpacing[0] > 0.0f) || !(b->raySpacing[1] > 0.0f))
            return fail(p, "Legacy CPB path requires explicit beam.raySpacing");
        float minGX = 1e30f, maxGX = -1e30f, minGY = 1e30f, maxGY = -1e30f;
        f3 size = mk3((float)(doseDims[0] > 1 ? doseDims[0] - 1 : 0) * doseRes_mm.x,
                      (float)(doseDims[1] > 1 ? doseDims[1] - 1 : 0) * doseRes_mm.y,
                      (float)(doseDims[2] > 1 ? doseDims[2] - 1 : 0) * doseRes_mm.z);
        for (int cx = 0; cx <= 1; ++cx)
            for (int cy = 0; cy <= 1; ++cy)
                for (int cz = 0; cz <= 1; ++cz) {
                    f3 w = add3(doseCorner_mm, mk3((float)cx * size.x, (float)cy * size.y, (float)cz * size.z));
                    f3 g = aff_apply(w2g, w);
                    minGX = fminf(minGX, g.x); maxGX = fmaxf(maxGX, g.x);
                    minGY = fminf(minGY, g.y); maxGY = fmaxf(maxGY, g.y);
                }
        for (int l = 0; l < b->numLayers; ++l)
            for (int s = 0; s < b->maxSubspots; ++s) {
                const float *ss = b->subspotData + ((size_t)l * b->maxSubspots + s) * 5;
                if (ss[2] > 0.0f) {
                    float sx = ss[0] / lenToMm, sy = ss[1] / lenToMm;
                    minGX = fminf(minGX, sx); maxGX = fmaxf(maxGX, sx);
                    minGY = fminf(minGY, sy); maxGY = fmaxf(maxGY, sy);
                }
            }
        float covX = (maxGX - minGX) / lenToMm, covY = (maxGY - minGY) / lenToMm;
        if (!(covX > 0.0f) || !(covY > 0.0f)) return fail(p, "Legacy CPB path produced non-positive coverage");
        if (b->cpbCoverage > 0.0f) covX = covY = b->cpbCoverage;
        p->cpbCorner[0] = minGX / lenToMm; p->cpbCorner[1] = minGY / lenToMm; p->cpbCorner[2] = 0.0f;
        p->cpbRes[0] = b->raySpacing[0]; p->cpbRes[1] = b->raySpacing[1]; p->cpbRes[2] = doseRes[2];
        int nx = (int)ceilf(covX / p->cpbRes[0]), ny = (int)ceilf(covY / p->cpbRes[1]);
        p->rayDims[0] = nx > 1 ? nx : 1;
        p->rayDims[1] = ny > 1 ? ny : 1;
        p->legacyPath = 1;
    }
    p->superpDims[0] = ((p->rayDims[0] + ORC_TILE_X - 1) / ORC_TILE_X) * ORC_TILE_X;
    p->superpDims[1] = ((p->rayDims[1] + ORC_TILE_Y - 1) / ORC_TILE_Y) * ORC_TILE_Y;

    /* index transforms (wrapper:3233-3248) */
    f3 ctRes_mm = mk3(ctRes[0] * lenToMm, ctRes[1] * lenToMm, ctRes[2] * lenToMm);
    f3 ctCorner_mm = mk3(ctCorner[0] * lenToMm, ctCorner[1] * lenToMm, ctCorner[2] * lenToMm);
    aff w2im, w2dose;
    w2im.m = mdiag3(1.0f / ctRes_mm.x, 1.0f / ctRes_mm.y, 1.0f / ctRes_mm.z);
    w2im.v = mk3(-ctCorner_mm.x / ctRes_mm.x, -ctCorner_mm.y / ctRes_mm.y, -ctCorner_mm.z / ctRes_mm.z);
    w2dose.m = mdiag3(1.0f / doseRes_mm.x, 1.0f / doseRes_mm.y, 1.0f / doseRes_mm.z);
    w2dose.v = mk3(-doseCorner_mm.x / doseRes_mm.x, -doseCorner_mm.y / doseRes_mm.y, -doseCorner_mm.z / doseRes_mm.z);
    aff g2im = aff_concat(g2w, w2im), g2dose = aff_concat(g2w, w2dose);
    m3_to9(g2im.m, p->g2im);
    p->g2imOff[0] = g2im.v.x; p->g2imOff[1] = g2im.v.y; p->g2imOff[2] = g2im.v.z;
    m3_to9(g2dose.m, p->g2dose);
    p->g2doseOff[0] = g2dose.v.x; p->g2doseOff[1] = g2dose.v.y; p->g2doseOff[2] = g2dose.v.z;

    /* source distances (wrapper:3250-3271) */
    float isx = 0.f, isy = 0.f;
    int inferred = infer_source_dist(b, bmX, bmY, bmZ, sad, &isx, &isy);
    float dxmm = b->sourceDist[0] * lenToMm, dymm = b->sourceDist[1] * lenToMm;
    int legacySad = fabsf(b->sourceDist[0] - b->sad) < 1.0e-3f && fabsf(b->sourceDist[1] - b->sad) < 1.0e-3f;
    if (inferred && (!(dxmm > 1e-6f) || legacySad)) dxmm = isx * lenToMm;
    if (inferred && (!(dymm > 1e-6f) || legacySad)) dymm = isy * lenToMm;
    if (!(dxmm > 1e-6f)) dxmm = sad_mm;
    if (!(dymm > 1e-6f)) dymm = sad_mm;
    p->sourceDist[0] = dxmm; p->sourceDist[1] = dymm;

    /* tracer extent (wrapper:3287-3370): gantry z range of the CT box, K steps of ctRes.z */
    float minGZ = 1e30f, maxGZ = -1e30f;
    f3 ctSize = mk3((float)(ctDims[0] > 1 ? ctDims[0] - 1 : 0) * ctRes_mm.x,
                    (float)(ctDims[1] > 1 ? ctDims[1] - 1 : 0) * ctRes_mm.y,
                    (float)(ctDims[2] > 1 ? ctDims[2] - 1 : 0) * ctRes_mm.z);
    for (int cx = 0; cx <= 1; ++cx)
        for (int cy = 0; cy <= 1; ++cy)
            for (int cz = 0; cz <= 1; ++cz) {
                f3 w = add3(ctCorner_mm, mk3((float)cx * ctSize.x, (float)cy * ctSize.y, (float)cz * ctSize.z));
                f3 g = aff_apply(w2g, w);
                minGZ = fminf(minGZ, g.z); maxGZ = fmaxf(maxGZ, g.z);
            }
    float stepLen = ctRes[2] * lenToMm;
    float startZ = maxGZ;
    p->weqVolume = b->weq && b->weqCount > 9;
    p->spotIndices = b->spotIndices != 0;
    if (p->weqVolume) {
        /* WEQ volume (wrapper:3305-3370): step length and count from the header; the start plane
         * moves to (mean source distance - depth_start) when that is within the volume's depth
         * span, or 50 mm, of the CT's */
        const float *h = b->weq;
        if (h[1] * lenToMm > 0.0f) stepLen = h[1] * lenToMm;
        float proj = h[0] * lenToMm;
        float aligned = 0.5f * (p->sourceDist[0] + p->sourceDist[1]) - proj;
        long nh = lroundf(h[2]);
        float span = fabsf(h[1] * lenToMm) * (float)(nh > 0 ? nh : 0);
        float allowed = span > 50.0f ? span : 50.0f;
        if (isfinite(aligned) && isfinite(proj) && proj > 0.0f && fabsf(aligned - maxGZ) <= allowed) startZ = aligned;
        p->steps = nh > 1 ? (int)nh : 1;
        int nx = (int)lroundf(h[8]), ny = (int)lroundf(h[5]);
        if (nx < 1) nx = 1;
        if (ny < 1) ny = 1;
        if (b->weqCount - 9 < (size_t)nx * ny * (size_t)p->steps) {
            snprintf(p->error, sizeof(p->error), "water_equivalence payload is smaller than header-declared nx*ny*nStep");
            return -1;
        }
        p->weqDims[0] = nx; p->weqDims[1] = ny; p->weqDims[2] = p->steps;
        p->weqStepMm = fabsf(h[1] * lenToMm);
        p->weqX0 = h[6] / lenToMm; p->weqY0 = h[3] / lenToMm;
        p->weqDx = h[7] / lenToMm; p->weqDy = h[4] / lenToMm;
        p->weqBody = h + 9;
    } else {
        int steps = (int)ceilf((maxGZ - minGZ) / fabsf(stepLen)) + 1;
        p->steps = steps > 1 ? steps : 1;
    }
    p->fanCorner[0] = p->cpbCorner[0] * lenToMm;
    p->fanCorner[1] = p->cpbCorner[1] * lenToMm;
    p->fanCorner[2] = startZ;
    p->fanDelta[0] = p->cpbRes[0] * lenToMm;
    p->fanDelta[1] = p->cpbRes[1] * lenToMm;
    p->fanDelta[2] = -fabsf(stepLen);

    /* frustum volume per step (wrapper:4276-4289), shared by all layers */
    {
        float dX = p->fanDelta[0], dY = p->fanDelta[1], dZ = p->fanDelta[2], cz = p->fanCorner[2];
        float dsx = p->sourceDist[0], dsy = p->sourceDist[1];
        float dxyz = fabsf(dX * dY * dZ);
        p->volConst = dxyz * (1.0f - cz / dsx - cz / dsy + (cz * cz + dZ * dZ / 12.0f) / (dsx * dsy));
        p->volLin = dxyz * dZ * (-1.0f / dsx - 1.0f / dsy + 2.0f * cz / (dsx * dsy));
        p->volSq = dxyz * dZ * dZ / (dsx * dsy);
    }
    plan_profile(b, en, p);
    p->ok = 1;
    return 0;
}

/* ------------------------------------------------------------------------------------
 * trace (bev_ray_tracing.cu:65-159 with ray_tracing.h:27-92 geometry)
 * ---------------------------------------------------------------------------------- */
typedef struct { double m[9], o[3], c[3], d[3], dist[2]; } fan_geom;

static void fan_geom_from(const float *m9, const float *off, const orc_plan *p, fan_geom *g) {
    for (int i = 0; i < 9; ++i) g->m[i] = m9[i];
    for (int i = 0; i < 3; ++i) { g->o[i] = off[i]; g->c[i] = p->fanCorner[i]; g->d[i] = p->fanDelta[i]; }
    g->dist[0] = p->sourceDist[0]; g->dist[1] = p->sourceDist[1];
}

/* Float3FromFanTransform::transformPoint in double (fan_transforms.h:39-49) */
static void fan_to_index(const fan_geom *g, double fi, double fj, double fk, double out[3]) {
    double x = g->c[0] + fi * g->d[0], y = g->c[1] + fj * g->d[1], z = g->c[2] + fk * g->d[2];
    x *= (1.0 - z / g->dist[0]);
    y *= (1.0 - z / g->dist[1]);
    for (int r = 0; r < 3; ++r) out[r] = g->m[3 * r] * x + g->m[3 * r + 1] * y + g->m[3 * r + 2] * z + g->o[r];
}

/* fillBevFromWeqVolumeKernel (wrapper:1923-1988): per ray, the nearest WEQ column (float index
 * arithmetic, rounded half away from zero as lroundf); WEPL(k) = sample k (the last value read
 * past the column or off the lattice), density = max(0, WEPL(k) - WEPL(k-1)) / step, entry / exit
 * steps from the first / last density > 1e-4 (+1; none: K and 0). */
static void weq_trace(const orc_plan *p, double *density, double *wepl, int *firstInside, int *firstOutside) {
    const int rx = p->rayDims[0], ry = p->rayDims[1], K = p->steps;
    const size_t plane = (size_t)rx * ry;
    const int nx = p->weqDims[0], ny = p->weqDims[1], ns = p->weqDims[2];
    const int active = K < ns ? K : ns;
    const double step = p->weqStepMm > 0.0f ? (double)p->weqStepMm : 1.0;
    for (int j = 0; j < ry; ++j) {
        for (int i = 0; i < rx; ++i) {
            const size_t r = (size_t)j * rx + i;
            float rayX = p->cpbCorner[0] + (float)i * p->cpbRes[0], rayY = p->cpbCorner[1] + (float)j * p->cpbRes[1];
            long qx = fabsf(p->weqDx) > 0.0f ? lroundf((rayX - p->weqX0) / p->weqDx) : -1;
            long qy = fabsf(p->weqDy) > 0.0f ? lroundf((rayY - p->weqY0) / p->weqDy) : -1;
            int in = qx >= 0 && qx < nx && qy >= 0 && qy < ny;
            const float *col = in ? p->weqBody + ((size_t)qx * ny + (size_t)qy) * ns : NULL;
            double prev = 0.0, last = 0.0;
            int first = -1, lastMat = -1;
            for (int k = 0; k < K; ++k) {
                double w = last;
                if (in && k < active) { w = (double)col[k]; last = w; }
                double d = (w - prev > 0.0 ? w - prev : 0.0) / step;
                wepl[(size_t)k * plane + r] = w;
                density[(size_t)k * plane + r] = d;
                if (d > 1.0e-4) { if (first < 0) first = k; lastMat = k; }
                prev = w;
            }
            firstInside[r] = first >= 0 ? first + 1 : K;
            firstOutside[r] = lastMat >= 0 ? lastMat + 1 : 0;
        }
    }
}

void orc_trace(const orc_plan *p, const float *ct, const int ctDims[3], const orc_energy *en,
               double *density, double *wepl, int *firstInside, int *firstOutside) {
    if (p->weqVolume) {
        weq_trace(p, density, wepl, firstInside, firstOutside);
        return;
    }
    const int rx = p->rayDims[0], ry = p->rayDims[1], K = p->steps;
    const size_t plane = (size_t)rx * ry;
    fan_geom g;
    fan_geom_from(p->g2im, p->g2imOff, p, &g);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ry; ++j) {
        for (int i = 0; i < rx; ++i) {
            double a[3], bpt[3];
            fan_to_index(&g, i, j, 0, a);
            fan_to_index(&g, i, j, 1, bpt);
            double inc[3] = {bpt[0] - a[0], bpt[1] - a[1], bpt[2] - a[2]};
            double ddx = (g.c[0] + i * g.d[0]) / g.dist[0], ddy = (g.c[1] + j * g.d[1]) / g.dist[1];
            double stepLen = fabs(g.d[2]) * sqrt(1.0 + ddx * ddx + ddy * ddy);
            double cumulSp = 0.0, cumulHu = 0.0;
            unsigned before = 0, lastIn = 0;
            int flag = 1;
            size_t idx = (size_t)j * rx + i;
            for (int k = 0; k < K; ++k) {
                double hu = orc_sample3d_border(ct, ctDims, a[0] + k * inc[0], a[1] + k * inc[1], a[2] + k * inc[2]);
                cumulHu += hu;
                density[idx] = lut1d_clamp(en->density, en->nDensity, hu * en->densityScale);
                cumulSp += stepLen * lut1d_clamp(en->sp, en->nSp, hu * en->spScale);
                if (flag && cumulHu < 150.0) before = (unsigned)k; else flag = 0;
                if (hu > 150.0) lastIn = (unsigned)k;
                wepl[idx] = cumulSp;
                idx += plane;
            }
            firstInside[(size_t)j * rx + i] = (int)before + 1;
            firstOutside[(size_t)j * rx + i] = (int)lastIn + 1;
        }
    }
}

void orc_beam_cutoffs(orc_plan *p, const orc_beam *b, const orc_energy *en, const double *wepl,
                      const int *fi, const int *fo, double *weplMin) {
    const int n = p->rayDims[0] * p->rayDims[1], K = p->steps;
    int mn = INT32_MAX, mx = INT32_MIN;
    for (int r = 0; r < n; ++r) { if (fi[r] < mn) mn = fi[r]; if (fo[r] > mx) mx = fo[r]; }
    p->beamFirstInside = mn;
    p->beamFirstOutside = mx;
    for (int k = 0; k < K; ++k) {
        double m = ORC_INF_F;
        for (int r = 0; r < n; ++r)
            if (fo[r] > 0 && wepl[(size_t)k * n + r] < m) m = wepl[(size_t)k * n + r];
        weplMin[k] = m;
    }
    float maxE = b->energies[0];
    for (int l = 1; l < b->numLayers; ++l) if (b->energies[l] > maxE) maxE = b->energies[l];
    float mi = orc_find_decimal_ordered(en->energies, en->nEnergies, maxE);
    p->maxPeakDepth = orc_vector_interpolate(en->peaks, en->nEnergies, mi) * p->energyDepthToMm;
    p->firstPastCutoffAll = orc_find_first_larger_ordered(weplMin, K, (double)(1.05f * p->maxPeakDepth));
    p->guaranteedPassive = p->firstPastCutoffAll < p->beamFirstOutside ? p->firstPastCutoffAll : p->beamFirstOutside;
}

/* ------------------------------------------------------------------------------------
 * spot -> ray convolution
 * ---------------------------------------------------------------------------------- */
static float sigma_to_mm(float lenToMm, float s) {
    if (lenToMm > 1.0f && s > 0.0f && s < 1.0f) return s * 10.0f;
    return s;
}

/* sigmaSqAirCoefs (fill_idd_and_sigma_params.cuh:76-78) */
static void air_coefs(float r0, float *a, float *b) {
    *a = 0.00270f / (r0 - 4.50f);
    *b = -4.39f / (r0 - 3.86f);
}

/* x/yConvGathResampGpu restated (convolution.cu:198-262): gather-resample with erf taps */
static void conv1d_gather(const double *in, int inW, int inStride, double *out, int outW, int outStride,
                          double sigma, double pixelSp, double inOutOffset, double inOutDelta) {
    const double sEff = sigma / pixelSp;
    const double rS = 1.0 / (sqrt(2.0) * sEff);
    const double reach = 3.0 * sEff + 0.5;
    for (int o = 0; o < outW; ++o) {
        double res = 0.0;
        int ci = (int)ceil(((double)o - reach - inOutOffset) / inOutDelta);
        double dist = ci * inOutDelta + inOutOffset - (double)o;
        while (dist < reach) {
            if (ci >= 0 && ci < inW)
                res += 0.5 * (erf((dist + 0.5) * rS) - erf((dist - 0.5) * rS)) * in[(size_t)ci * inStride];
            ++ci;
            dist = ci * inOutDelta + inOutOffset - (double)o;
        }
        out[(size_t)o * outStride] = res;
    }
}

static void conv_plane(const double *spot, int sx, int sy, double *rays, int rx, int ry,
                       double sigX, double sigY, const orc_plan *p, float offX, float offY,
                       double pxMultX, double pxMultY, double scale, int accumulate) {
    double *interm = (double *)malloc(sizeof(double) * (size_t)rx * sy);
    double *tmp = (double *)malloc(sizeof(double) * (size_t)rx * ry);
    /* performExactRTDConvolution2D (convolution.cu:477-513) */
    float dX = p->spotDelta[0] / p->cpbRes[0], dY = p->spotDelta[1] / p->cpbRes[1];
    float oX = (offX - p->cpbCorner[0]) / p->cpbRes[0], oY = (offY - p->cpbCorner[1]) / p->cpbRes[1];
    for (int y = 0; y < sy; ++y)
        conv1d_gather(spot + (size_t)y * sx, sx, 1, interm + (size_t)y * rx, rx, 1, sigX,
                      (double)p->cpbRes[0] * pxMultX, oX, dX);
    for (int x = 0; x < rx; ++x)
        conv1d_gather(interm + x, sy, rx, tmp + x, ry, rx, sigY, (double)p->cpbRes[1] * pxMultY, oY, dY);
    for (size_t i = 0; i < (size_t)rx * ry; ++i) rays[i] = (accumulate ? rays[i] : 0.0) + scale * tmp[i];
    free(interm); free(tmp);
}

void orc_conv_lattice(orc_plan *p, const orc_beam *b, const orc_energy *en, double *rayWeights) {
    const int rx = p->rayDims[0], ry = p->rayDims[1];
    const int sx = p->spotDims[0], sy = p->spotDims[1];
    float *dense = build_lattice(b, p);
    const float entryZ = (float)p->beamFirstInside * p->fanDelta[2] + p->fanCorner[2];
    const double pxX = 1.0f - entryZ / p->sourceDist[0], pxY = 1.0f - entryZ / p->sourceDist[1];
    double *spot = (double *)malloc(sizeof(double) * (size_t)sx * sy);
    for (int l = 0; l < b->numLayers; ++l) {
        float eIdx = orc_find_decimal_ordered(en->energies, en->nEnergies, b->energies[l]);
        float peak = orc_vector_interpolate(en->peaks, en->nEnergies, eIdx) * p->energyDepthToMm;
        float ca, cb;
        air_coefs(peak, &ca, &cb);
        p->convPeakDepth[l] = peak;
        for (size_t i = 0; i < (size_t)sx * sy; ++i) spot[i] = dense[(size_t)l * sx * sy + i];
        double *out = rayWeights + (size_t)l * rx * ry;
        int done = 0;
        if (b->useSubspotConvolution && b->maxSubspots > 0) {
            /* per-subspot branch (wrapper:3770-3822) */
            for (int s = 0; s < b->maxSubspots; ++s) {
                const float *ss = b->subspotData + ((size_t)l * b->maxSubspots + s) * 5;
                if (!(ss[2] > 1.0e-8f)) continue;
                float sgx = sigma_to_mm(p->lenToMm, ss[3]), sgy = sigma_to_mm(p->lenToMm, ss[4]);
                if (!(sgx > 1.0e-8f) || !(sgy > 1.0e-8f)) continue;
                float vx = ca * entryZ * entryZ + cb * entryZ + sgx * sgx;
                float vy = ca * entryZ * entryZ + cb * entryZ + sgy * sgy;
                double ex = sqrt((double)(vx > 0.f ? vx : 0.f)), ey = sqrt((double)(vy > 0.f ? vy : 0.f));
                if (b->nuclear == 1) { ex *= 0.97; ey *= 0.97; } /* GAUSS_FIT only, wrapper:3788-3795 */
                conv_plane(spot, sx, sy, out, rx, ry, ex, ey, p,
                           p->spotOffset[0] + ss[0] / p->lenToMm, p->spotOffset[1] + ss[1] / p->lenToMm,
                           pxX, pxY, (double)ss[2], done);
                done = 1;
            }
        }
        if (!done) {
            float sgx = sigma_to_mm(p->lenToMm, b->spotSigmas[2 * l]);
            float sgy = sigma_to_mm(p->lenToMm, b->spotSigmas[2 * l + 1]);
            float vx = ca * entryZ * entryZ + cb * entryZ + sgx * sgx;
            float vy = ca * entryZ * entryZ + cb * entryZ + sgy * sgy;
            double ex = sqrt((double)(vx > 0.f ? vx : 0.f)), ey = sqrt((double)(vy > 0.f ? vy : 0.f));
            if (b->nuclear == 1) { ex *= 0.97; ey *= 0.97; } /* GAUSS_FIT only, wrapper:3788-3795 */
            p->entrySigma[l][0] = (float)ex; p->entrySigma[l][1] = (float)ey;
            conv_plane(spot, sx, sy, out, rx, ry, ex, ey, p, p->spotOffset[0], p->spotOffset[1], pxX, pxY, 1.0, 0);
        }
    }
    free(spot);
    free(dense);
}

void orc_subspot_to_cpb(const float *ssd, int L, int S, const float cc[3], const float cr[3], const int cd[3],
                        const float bd[3], const float bx[3], const float by[3], const float src[3], float sad,
                        double *cpb) {
    memset(cpb, 0, sizeof(double) * (size_t)cd[0] * cd[1] * cd[2]);
    for (int l = 0; l < L && l < cd[2]; ++l) {
        for (int s = 0; s < S; ++s) {
            const float *q = ssd + ((size_t)l * S + s) * 5;
            float dX = q[0], dY = q[1], w = q[2], sX = q[3], sY = q[4];
            /* SubspotInfo (common.cuh:172-217): position on the reference plane */
            f3 pos = add3(add3(add3(mk3(src[0], src[1], src[2]), mul3s(mk3(bd[0], bd[1], bd[2]), sad)),
                               mul3s(mk3(bx[0], bx[1], bx[2]), dX)), mul3s(mk3(by[0], by[1], by[2]), dY));
            f3 sd = sub3(pos, mk3(src[0], src[1], src[2]));
            if (!(sqrtf(dot3(sd, sd)) > 1e-6f)) continue;
            if (!((w > 1e-6f) && (sX > 1e-6f) && (sY > 1e-6f))) continue;
            /* calculateSubspotRangeKernel (convolution.cu:45-88) */
            float minX = pos.x - 3.0f * sX, maxX = pos.x + 3.0f * sX;
            float minY = pos.y - 3.0f * sY, maxY = pos.y + 3.0f * sY;
            int x0 = (int)floorf((minX - cc[0]) / cr[0]); if (x0 < 0) x0 = 0;
            int x1 = (int)ceilf((maxX - cc[0]) / cr[0]); if (x1 > cd[0] - 1) x1 = cd[0] - 1;
            int y0 = (int)floorf((minY - cc[1]) / cr[1]); if (y0 < 0) y0 = 0;
            int y1 = (int)ceilf((maxY - cc[1]) / cr[1]); if (y1 > cd[1] - 1) y1 = cd[1] - 1;
            const double r2x = 1.0 / (sqrt(2.0) * sX), r2y = 1.0 / (sqrt(2.0) * sY);
            for (int cy = y0; cy <= y1; ++cy)
                for (int cx = x0; cx <= x1; ++cx) {
                    double px = (double)cc[0] + (cx + 0.5) * cr[0], py = (double)cc[1] + (cy + 0.5) * cr[1];
                    double ddx = px - pos.x, ddy = py - pos.y;
                    double wx = erf((ddx + 0.5 * cr[0]) * r2x) - erf((ddx - 0.5 * cr[0]) * r2x);
                    double wy = erf((ddy + 0.5 * cr[1]) * r2y) - erf((ddy - 0.5 * cr[1]) * r2y);
                    cpb[((size_t)l * cd[1] + cy) * cd[0] + cx] += 0.25 * wx * wy * w;
                }
        }
    }
}

void orc_cpb_to_rays(const double *cpb, const int cd[3], const float cc[3], const float cr[3],
                     double *rays, const int rd[2], const float rc[3], const float rr[3], int layer) {
    for (int ry = 0; ry < rd[1]; ++ry)
        for (int rx = 0; rx < rd[0]; ++rx) {
            double x0 = (double)rc[0] + rx * (double)rr[0], x1 = x0 + rr[0];
            double y0 = (double)rc[1] + ry * (double)rr[1], y1 = y0 + rr[1];
            double acc = 0.0;
            double area = (double)cr[0] * cr[1];
            if (layer >= 0 && layer < cd[2] && area > 0.0) {
                int cx0 = (int)floor((x0 - cc[0]) / cr[0]), cx1 = (int)ceil((x1 - cc[0]) / cr[0]) - 1;
                int cy0 = (int)floor((y0 - cc[1]) / cr[1]), cy1 = (int)ceil((y1 - cc[1]) / cr[1]) - 1;
                if (cx0 < 0) cx0 = 0; if (cy0 < 0) cy0 = 0;
                if (cx1 >= cd[0]) cx1 = cd[0] - 1; if (cy1 >= cd[1]) cy1 = cd[1] - 1;
                for (int cy = cy0; cy <= cy1; ++cy) {
                    double c0 = (double)cc[1] + cy * (double)cr[1], c1 = c0 + cr[1];
                    double oy = fmax(0.0, fmin(y1, c1) - fmax(y0, c0));
                    if (oy <= 0.0) continue;
                    for (int cx = cx0; cx <= cx1; ++cx) {
                        double d0 = (double)cc[0] + cx * (double)cr[0], d1 = d0 + cr[0];
                        double ox = fmax(0.0, fmin(x1, d1) - fmax(x0, d0));
                        if (ox <= 0.0) continue;
                        acc += cpb[((size_t)layer * cd[1] + cy) * cd[0] + cx] * (ox * oy) / area;
                    }
                }
            }
            rays[(size_t)ry * rd[0] + rx] = acc;
        }
}

/* ------------------------------------------------------------------------------------
 * per-layer parameters (wrapper:3917-4386)
 * ---------------------------------------------------------------------------------- */
static float layer_energy_index(const orc_energy *en, float energy) {
    const int n = en->nEnergies;
    if (n <= 1) return 0.0f;
    int asc = en->energies[0] < en->energies[n - 1];
    int lo = asc ? 0 : n - 1, hi = asc ? n - 1 : 0;
    float eLo = en->energies[lo], eHi = en->energies[hi];
    if (energy <= eLo) return (float)lo;
    if (energy >= eHi) return (float)hi;
    int fl = 0;
    float e0 = en->energies[0], e1 = en->energies[0];
    for (int i = 0; i < n - 1; ++i) {
        float a = en->energies[i], b = en->energies[i + 1];
        float mn = a < b ? a : b, mx = a < b ? b : a;
        if (energy >= mn && energy <= mx) { fl = i; e0 = a; e1 = b; break; }
    }
    float den = e1 - e0;
    float corr = (den != 0.0f) ? ((energy - e0) / den) : 0.0f;
    if (!isfinite(corr)) corr = 0.0f;
    corr = corr < 0.0f ? 0.0f : (corr > 1.0f ? 1.0f : corr);
    return (float)fl + corr;
}

static float lerp_table(const float *v, int n, float idx) {
    if (idx <= 0.0f) return v[0];
    if (idx >= (float)(n - 1)) return v[n - 1];
    float ip;
    float dec = modff(idx, &ip);
    int f = (int)ip;
    return v[f] + (v[f + 1] - v[f]) * dec;
}

void orc_layer_params(orc_plan *p, const orc_beam *b, const orc_energy *en, const double *weplMin) {
    for (int l = 0; l < b->numLayers; ++l) {
        float eIdx = layer_energy_index(en, b->energies[l]);
        float sf = lerp_table(en->scales, en->nEnergies, eIdx) / p->energyDepthToMm;
        float pk = lerp_table(en->peaks, en->nEnergies, eIdx) * p->energyDepthToMm;
        p->energyIdx[l] = eIdx;
        p->scaleFact[l] = sf;
        p->peakDepth[l] = pk;
        p->entrySpotSigmaX[l] = sigma_to_mm(p->lenToMm, b->spotSigmas[2 * l]);
        float a, c;
        air_coefs(pk, &a, &c);
        /* initStepAndAirDiv (fill_idd_and_sigma_params.cuh:66-73) */
        p->sigmaSqAirQuad[l] = a * 1.0f * p->fanDelta[2] * p->fanDelta[2];
        p->sigmaSqAirLin[l] = 2.0f * a * 1.0f * p->fanDelta[2] * p->fanCorner[2] + c * p->fanDelta[2];
        int local = orc_find_first_larger_ordered(weplMin, p->steps, (double)(1.05f * pk));
        p->afterLast[l] = local < p->guaranteedPassive ? local : p->guaranteedPassive;
    }
}

/* ------------------------------------------------------------------------------------
 * IDD and sigma (idd_sigma.cu:46-259)
 * ---------------------------------------------------------------------------------- */
/* Scattering constants of the halo models (idd_sigma.cu:108-121): eRefSq, sigmaDelta. */
static void halo_model_consts(int model, double *eRefSq, double *sigmaDelta, int *suppressDip) {
    *eRefSq = 198.81; *sigmaDelta = 0.21; *suppressDip = 0;
    if (model == 1) { *eRefSq = 169.00; *sigmaDelta = 0.06; *suppressDip = 1; } /* GAUSS_FIT */
    if (model == 2) { *eRefSq = 190.44; *sigmaDelta = 0.0; }                     /* SOUKUP */
    if (model == 3) { *eRefSq = 216.09; *sigmaDelta = 0.08; }                    /* FLUKA */
}

void orc_fill_halo(const orc_plan *p, int l, const orc_energy *en, const double *density, const double *wepl,
                   const double *rayWeights, const int *fi, const int *fo, double *idd, double *rsigma,
                   int *firstPassive, int model, const int *rayToNuc, const double *nucW, double *nucIdd,
                   double *nucRs) {
    const int rx = p->rayDims[0], ry = p->rayDims[1];
    const size_t plane = (size_t)rx * ry;
    const size_t nucPlane = (size_t)p->nucDims[0] * p->nucDims[1];
    const unsigned first = (unsigned)p->beamFirstInside, pAfter = (unsigned)p->afterLast[l];
    const double peak = p->peakDepth[l], rangeStop = 1.05f * p->peakDepth[l];
    const double scale = p->scaleFact[l], eIdx = p->energyIdx[l];
    const double stepLength = fabs((double)p->fanDelta[2]);
    const double quad = p->sigmaSqAirQuad[l], lin = p->sigmaSqAirLin[l];
    double eRefSq, sigmaDelta;
    int suppressDip;
    halo_model_consts(model, &eRefSq, &sigmaDelta, &suppressDip);
    /* the table set's ion (proton defaults: idd_sigma.cu:172-181) */
    const double rc = en->ionRangeCoef > 0.0f ? (double)en->ionRangeCoef : 8.639415;
    const double rq = en->ionRangeExp > 0.0f ? (double)en->ionRangeExp : 0.5649718;
    const double M = en->ionRestMass > 0.0f ? (double)en->ionRestMass : 938.3;
    if (en->ionScatterScale > 0.0f) eRefSq *= (double)en->ionScatterScale;
    const int halo = model != 0 && rayToNuc && nucW && nucIdd && nucRs && en->nucWeight && en->nucSqSigma &&
                     nucPlane > 0 && p->spotDist > 0.0f;
    /* entry sigma^2 of the halo: sigmaSqAirCoefs(peak) at the entry plane + spot sigma_x^2
     * (wrapper:4324-4329), float as on the host */
    float entrySigmaSq = 0.0f;
    if (halo) {
        float a, c;
        air_coefs(p->peakDepth[l], &a, &c);
        const float entryZ = (float)p->beamFirstInside * p->fanDelta[2] + p->fanCorner[2];
        const float sgx = p->entrySpotSigmaX[l];
        entrySigmaSq = a * entryZ * entryZ + c * entryZ + sgx * sgx;
    }
    const double spotDist = p->spotDist;
#pragma omp parallel for schedule(static)
    for (int r = 0; r < (int)plane; ++r) {
        const unsigned firstIn = (unsigned)fi[r];
        int al = fo[r] < (int)pAfter ? fo[r] : (int)pAfter;
        unsigned afterLast = (unsigned)al;
        const double w = rayWeights[r];
        int live = 1;
        if (w < 1e-6 || afterLast < first) { live = 0; afterLast = 0; }
        const int cell = halo ? rayToNuc[r] : -1;
        const double nucRayW = cell >= 0 ? nucW[cell] : 0.0;
        double cumulSp = 0, cumulSpOld = 0, cumulDose = 0, cumulDoseOld = 0;
        double incScat = 0, incincScat = 0;
        double incDiv = lin + (2.0 * (double)first - 1.0) * quad;
        double sigmaSq = -incDiv;
        for (unsigned k = first; k < pAfter; ++k) {
            double res = 0.0, rs = INFINITY, nres = 0.0, nrs = INFINITY;
            const size_t idx = (size_t)k * plane + r;
            if (live) {
                cumulSp = wepl[idx];
                cumulDose = lut2d_clamp(en->cidd, en->nSamples, en->nEnergies, cumulSp * scale, eIdx);
                const double rho = density[idx];
                if (cumulSp < peak) {
                    double resE = rc * pow(peak - 0.5 * (cumulSp + cumulSpOld), rq);
                    double betaP = resE + M - M * M / (resE + M);
                    double rRl = rho * lut1d_clamp(en->rrl, en->nRrl, rho * en->rrlScale);
                    double thetaSq = eRefSq / (betaP * betaP) * stepLength * rRl;
                    sigmaSq += incScat + incDiv;
                    incincScat += 2.0 * thetaSq * stepLength * stepLength;
                    incScat += incincScat;
                    incDiv += 2.0 * quad;
                } else if (!suppressDip) {
                    sigmaSq -= 1.5 * (incScat + incDiv) * rho;
                }
                double zk = (double)p->fanCorner[2] + (double)k * p->fanDelta[2];
                double vwx = p->fanDelta[0] * (1.0 - zk / p->sourceDist[0]);
                double vwy = p->fanDelta[1] * (1.0 - zk / p->sourceDist[1]);
                rs = 0.5 * (vwx + vwy) / (sqrt(2.0) * (sqrt(sigmaSq) + sigmaDelta));
                double vol = (double)p->volConst + (double)k * p->volLin + (double)k * (double)k * p->volSq;
                double mass = rho * vol;
                const double depthMid = 0.5 * (cumulSp + cumulSpOld) * scale;
                if (mass > 1e-2) {
                    if (halo) {
                        const double nw = lut2d_clamp(en->nucWeight, en->nSamples, en->nEnergies, depthMid, eIdx);
                        res = (1.0 - nw) * w * (cumulDose - cumulDoseOld) / mass;
                        if (cell >= 0) nres = nw * nucRayW * (cumulDose - cumulDoseOld) / (mass * spotDist * spotDist);
                    } else {
                        res = w * (cumulDose - cumulDoseOld) / mass;
                    }
                }
                if (cell >= 0) {
                    const double s2 = lut2d_clamp(en->nucSqSigma, en->nSamples, en->nEnergies, depthMid, eIdx);
                    nrs = 0.5 * spotDist * (vwx + vwy) / (sqrt(2.0) * sqrt(sigmaSq + s2 + (double)entrySigmaSq));
                }
                cumulSpOld = cumulSp;
                cumulDoseOld = cumulDose;
                if (cumulSp > rangeStop || k == afterLast) { live = 0; afterLast = k; }
            }
            if (!live || (int)k < (int)firstIn - 1) { res = 0.0; rs = INFINITY; nres = 0.0; nrs = INFINITY; }
            idd[idx] = res;
            rsigma[idx] = rs;
            if (cell >= 0) {
                nucIdd[(size_t)k * nucPlane + cell] = nres;
                nucRs[(size_t)k * nucPlane + cell] = nrs;
            }
        }
        firstPassive[r] = (int)afterLast;
    }
}

void orc_fill(const orc_plan *p, int l, const orc_energy *en, const double *density, const double *wepl,
              const double *rayWeights, const int *fi, const int *fo, double *idd, double *rsigma,
              int *firstPassive) {
    orc_fill_halo(p, l, en, density, wepl, rayWeights, fi, fo, idd, rsigma, firstPassive, 0, NULL, NULL, NULL, NULL);
}

/* One axis of the physical PB lattice (inferAxis, wrapper:1555-1634). The distinct values are
 * the first seen of each 0.05 mm cluster, sorted. The step is the median gap; each gap must be
 * k steps with k in [1, 32] and |gap/step - k| <= 1/4. Every value must then lie within a
 * quarter cell of a point of the count = round(extent / step) + 1 lattice. Returns 1 on success. */
static int pb_axis_infer(const float *v, int n, float *delta, float *offset, int *count) {
    if (n <= 0) return 0;
    float *u = (float *)malloc(sizeof(float) * (size_t)n);
    int nu = 0;
    for (int i = 0; i < n; ++i) {
        int j = 0;
        while (j < nu && !(fabsf(v[i] - u[j]) <= 0.05f)) ++j;
        if (j == nu) u[nu++] = v[i];
    }
    qsort(u, (size_t)nu, sizeof(float), cmp_float);
    *offset = u[0];
    if (nu == 1) { *delta = 1.0f; *count = 1; free(u); return 1; }
    const size_t ng = (size_t)(nu > 1 ? nu - 1 : 1);  /* nu >= 2 here */
    float *g = (float *)malloc(sizeof(float) * ng);
    float *gs = (float *)malloc(sizeof(float) * ng);
    for (int i = 0; i + 1 < nu; ++i) g[i] = gs[i] = u[i + 1] - u[i];
    qsort(gs, (size_t)(nu - 1), sizeof(float), cmp_float);
    const float med = gs[(nu - 1) / 2];
    int ok = med > 0.0f;
    for (int i = 0; ok && i + 1 < nu; ++i) {
        const float r = g[i] / med, k = roundf(r);
        if (k < 1.0f || k > 32.0f || fabsf(r - k) > 0.25f) ok = 0;
    }
    if (ok) {
        *delta = med;
        *count = (int)lroundf((u[nu - 1] - u[0]) / med) + 1;
        if (*count < 1) *count = 1;
        for (int i = 0; ok && i < n; ++i) {
            const float f = (v[i] - *offset) / *delta, k = roundf(f);
            if (k < 0.0f || k > (float)(*count - 1) || fabsf(f - k) > 0.25f) ok = 0;
        }
    }
    free(u); free(g); free(gs);
    return ok;
}

/* The physical PB lattice of index-coded spots (buildPhysicalPBLatticeView, wrapper:1478-1820):
 * texel coordinates decoded to mm, both axes inferred over all spots; failing that, over the
 * first layer with the most spots (>= 3), with origin and count from all spots' extent; failing
 * that, the WEQ header's lattice, where spots off it are clamped to the edge. Weights go to the
 * nearest cell. Returns the dense [L][ny][nx] lattice and fills p->nuc{Raw,Offset,Delta}. */
static float *pb_lattice_view(const orc_beam *b, orc_plan *p) {
    const float *h = b->weq;
    if (!h || b->weqCount < 9) return NULL;
    const int n = b->numSpots;
    float *x = (float *)malloc(sizeof(float) * (size_t)n), *y = (float *)malloc(sizeof(float) * (size_t)n);
    for (int i = 0; i < n; ++i) {
        x[i] = h[6] + (b->spotPositions[2 * i] - 0.5f) * h[7];
        y[i] = h[3] + (b->spotPositions[2 * i + 1] - 0.5f) * h[4];
    }
    float dx = 0, dy = 0, ox = 0, oy = 0;
    int nx = 0, ny = 0, fallback = 0;
    if (!(pb_axis_infer(x, n, &dx, &ox, &nx) && pb_axis_infer(y, n, &dy, &oy, &ny))) {
        int best = -1, bestN = 0, first = 0, bestFirst = 0;
        for (int l = 0; l < b->numLayers; ++l) {
            if (b->layerSpotCounts[l] > bestN) { bestN = b->layerSpotCounts[l]; best = l; bestFirst = first; }
            first += b->layerSpotCounts[l];
        }
        int ok = 0;
        float ldx, ldy, lox, loy;
        int lnx, lny;
        if (best >= 0 && bestN >= 3 && pb_axis_infer(x + bestFirst, bestN, &ldx, &lox, &lnx) &&
            pb_axis_infer(y + bestFirst, bestN, &ldy, &loy, &lny)) {
            float x0 = x[0], x1 = x[0], y0 = y[0], y1 = y[0];
            for (int i = 1; i < n; ++i) {
                x0 = fminf(x0, x[i]); x1 = fmaxf(x1, x[i]);
                y0 = fminf(y0, y[i]); y1 = fmaxf(y1, y[i]);
            }
            dx = ldx; dy = ldy; ox = x0; oy = y0;
            nx = (int)lroundf((x1 - x0) / dx) + 1;
            ny = (int)lroundf((y1 - y0) / dy) + 1;
            if (nx < 1) nx = 1;
            if (ny < 1) ny = 1;
            ok = 1;
        }
        if (!ok) {
            nx = (int)lroundf(h[8]); if (nx < 1) nx = 1;
            ny = (int)lroundf(h[5]); if (ny < 1) ny = 1;
            dx = h[7]; dy = h[4]; ox = h[6]; oy = h[3];
            fallback = 1;
        }
    }
    float *dense = NULL;
    if (dx > 0.0f && dy > 0.0f && nx > 0 && ny > 0) {
        dense = (float *)calloc((size_t)nx * ny * b->numLayers, sizeof(float));
        int s = 0;
        for (int l = 0; l < b->numLayers && dense; ++l)
            for (int i = 0; i < b->layerSpotCounts[l]; ++i, ++s) {
                const float w = b->spotWeights[s];
                if (!isfinite(w) || w < 0.0f) { free(dense); dense = NULL; break; }
                if (w == 0.0f) continue;
                int ix = (int)lroundf((x[s] - ox) / dx), iy = (int)lroundf((y[s] - oy) / dy);
                if (ix < 0 || ix >= nx || iy < 0 || iy >= ny) {
                    if (!fallback) { free(dense); dense = NULL; break; }
                    ix = ix < 0 ? 0 : (ix >= nx ? nx - 1 : ix);
                    iy = iy < 0 ? 0 : (iy >= ny ? ny - 1 : iy);
                }
                dense[((size_t)l * ny + iy) * nx + ix] += w;
            }
    }
    free(x); free(y);
    if (dense) {
        p->nucRaw[0] = nx; p->nucRaw[1] = ny;
        p->nucOffset[0] = ox; p->nucOffset[1] = oy;
        p->nucDelta[0] = dx; p->nucDelta[1] = dy;
        p->nucFallback = fallback;
    }
    return dense;
}

int orc_halo_plan(orc_plan *p, const orc_beam *b, int *rayToNuc, double *nucWeights) {
    float *dense;
    if (b->spotIndices) {
        dense = pb_lattice_view(b, p);
        if (!dense) {
            snprintf(p->error, sizeof(p->error), "nuclear_correction=true requires a physical PB lattice view");
            return -1;
        }
    } else {
        dense = build_lattice(b, p);
        if (!dense) {
            snprintf(p->error, sizeof(p->error), "nuclear_correction=true requires raw spot lattice inputs");
            return -1;
        }
        p->nucRaw[0] = p->spotDims[0]; p->nucRaw[1] = p->spotDims[1];
        p->nucOffset[0] = p->spotOffset[0]; p->nucOffset[1] = p->spotOffset[1];
        p->nucDelta[0] = p->spotDelta[0]; p->nucDelta[1] = p->spotDelta[1];
        p->nucFallback = 0;
    }
    const int rawX = p->nucRaw[0], rawY = p->nucRaw[1];
    const int nucX = (rawX + ORC_TILE_X - 1) / ORC_TILE_X * ORC_TILE_X;
    const int nucY = (rawY + ORC_TILE_Y - 1) / ORC_TILE_Y * ORC_TILE_Y;
    p->nucDims[0] = nucX;
    p->nucDims[1] = nucY;
    const int rx = p->rayDims[0], ry = p->rayDims[1];
    for (int i = 0; i < rx * ry; ++i) rayToNuc[i] = -1;
    for (int j = 0; j < rawY; ++j) {
        const float gy = p->nucOffset[1] + (float)j * p->nucDelta[1];
        const int rayY = (int)lroundf((gy - p->cpbCorner[1]) / p->cpbRes[1]);
        if (rayY < 0 || rayY >= ry) {
            snprintf(p->error, sizeof(p->error), "dense halo lattice row mapped outside active ray grid");
            free(dense);
            return -1;
        }
        for (int i = 0; i < rawX; ++i) {
            const float gx = p->nucOffset[0] + (float)i * p->nucDelta[0];
            const int rayX = (int)lroundf((gx - p->cpbCorner[0]) / p->cpbRes[0]);
            if (rayX < 0 || rayX >= rx) {
                snprintf(p->error, sizeof(p->error), "dense halo lattice column mapped outside active ray grid");
                free(dense);
                return -1;
            }
            rayToNuc[rayY * rx + rayX] = j * nucX + i;
        }
    }
    if (nucWeights) {
        for (int l = 0; l < b->numLayers; ++l)
            for (int j = 0; j < nucY; ++j)
                for (int i = 0; i < nucX; ++i)
                    nucWeights[((size_t)l * nucY + j) * nucX + i] =
                        (i < rawX && j < rawY) ? dense[((size_t)l * rawY + j) * rawX + i] : 0.0;
    }
    free(dense);
    /* spot spacing: the beam's spotDelta when set (> 1.5 rays), else the halo lattice's
     * (wrapper:4149-4171) */
    const int beamDelta = b->spotDelta[0] > b->raySpacing[0] * 1.5f;
    const float spacing = beamDelta ? b->spotDelta[0] : p->nucDelta[0];
    if (b->raySpacing[0] > 0.0f && spacing / b->raySpacing[0] < 1.5f) {
        snprintf(p->error, sizeof(p->error), "nuclear superposition gate: spotDistInRays < 1.5");
        return -1;
    }
    p->spotDist = fabsf(p->fanDelta[0]) > 1e-6f ? fabsf(spacing / p->fanDelta[0]) : 1.0f;
    return 0;
}

/* ------------------------------------------------------------------------------------
 * superposition (complete_superposition.cu:197-544)
 * ---------------------------------------------------------------------------------- */
/* Superposition of steps [k0, k1) of a [K][sy][sx] field into a BEV [bevZ][sy+64][sx+64].
 * Radius histogram into hist[34] and batching into batch[33]. overflowIsError: the halo
 * grid's policy (complete_superposition.cu:558-620: a tile past the largest radius is an
 * error); the primary grid clamps it into the largest radius (413-463). Returns -1 on an
 * overflow that is an error, else 0. */
static int superpose_grid(int k0, int k1, int sx, int sy, const double *idd, const double *rsigma, double *bev,
                          int *hist, int *batch, int overflowIsError) {
    const int ntx = sx / ORC_TILE_X, nty = sy / ORC_TILE_Y;
    const int bx = sx + 2 * ORC_MAX_RAD, by = sy + 2 * ORC_MAX_RAD;
    memset(hist, 0, sizeof(int) * (ORC_MAX_RAD + 2));
    memset(batch, 0, sizeof(int) * (ORC_MAX_RAD + 1));
    if (k1 <= k0) return 0;
    const int nk = k1 - k0;
    int *rad = (int *)malloc(sizeof(int) * (size_t)nk * ntx * nty);
    for (int k = k0; k < k1; ++k)
        for (int ty = 0; ty < nty; ++ty)
            for (int tx = 0; tx < ntx; ++tx) {
                double m = NAN;
                for (int y = 0; y < ORC_TILE_Y; ++y)
                    for (int x = 0; x < ORC_TILE_X; ++x)
                        m = fmin(m, rsigma[((size_t)k * sy + ty * ORC_TILE_Y + y) * sx + tx * ORC_TILE_X + x]);
                int r = orc_tile_radius_f32((float)m);
                rad[((size_t)(k - k0) * nty + ty) * ntx + tx] = r;
                hist[r]++;
            }
    if (overflowIsError && hist[ORC_MAX_RAD + 1] > 0) {
        free(rad);
        return -1;
    }
    int effRad[ORC_MAX_RAD + 2];
    orc_batch_bins(hist, batch, effRad);
    for (int k = k0; k < k1; ++k) {
        double *plane = bev + (size_t)k * bx * by;
        /* a tile's block reaches at most 32 columns past its edges, so tile columns three
         * apart never overlap: three colour passes keep the per-pixel summation order fixed */
        for (int pass = 0; pass < 3; ++pass) {
#pragma omp parallel for schedule(dynamic)
            for (int tx = pass; tx < ntx; tx += 3) {
                double e[ORC_MAX_RAD + 1];
                for (int ty = 0; ty < nty; ++ty) {
                    const int R = effRad[rad[((size_t)(k - k0) * nty + ty) * ntx + tx]];
                    for (int y = 0; y < ORC_TILE_Y; ++y)
                        for (int x = 0; x < ORC_TILE_X; ++x) {
                            const size_t ii = ((size_t)k * sy + ty * ORC_TILE_Y + y) * sx + tx * ORC_TILE_X + x;
                            const double d = idd[ii];
                            if (d == 0.0) continue;
                            const double r = rsigma[ii];
                            double en = erf(r * 0.5), eo = -en;
                            for (int i = 0; i <= R; ++i) {
                                e[i] = 0.5 * (en - eo);
                                eo = en;
                                en = erf(r * ((double)i + 1.5));
                            }
                            const int X = tx * ORC_TILE_X + x + ORC_MAX_RAD;
                            const int Y = ty * ORC_TILE_Y + y + ORC_MAX_RAD;
                            for (int j = -R; j <= R; ++j) {
                                double *row = plane + (size_t)(Y + j) * bx + X;
                                const double dj = d * e[j < 0 ? -j : j];
                                for (int i = -R; i <= R; ++i) row[i] += dj * e[i < 0 ? -i : i];
                            }
                        }
                }
            }
        }
    }
    free(rad);
    return 0;
}

void orc_superpose(orc_plan *p, int l, const double *idd, const double *rsigma, double *bev) {
    (void)superpose_grid(p->beamFirstInside, p->layerPassive[l], p->superpDims[0], p->superpDims[1], idd, rsigma, bev,
                         p->radHist[l], p->batchRad[l], 0);
}

/* ------------------------------------------------------------------------------------
 * BEV -> dose transfer (wrapper:5343-5505, transfer_param_struct_div3*, prim_transf_kernel.cu)
 * ---------------------------------------------------------------------------------- */
/* A fan grid: cell (i, j) of step k is the fan point (fc + (i, j) * fd, fanCorner.z +
 * k * fanDelta.z). The primary grid's cells are the rays; the halo grid's cells are the
 * spots (nucIdxToFan, wrapper:5527-5535). */
typedef struct { int sx, sy; float fc[2], fd[2]; } fan_grid;

static fan_grid primary_grid(const orc_plan *p) {
    fan_grid g = {p->superpDims[0], p->superpDims[1], {p->fanCorner[0], p->fanCorner[1]},
                  {p->fanDelta[0], p->fanDelta[1]}};
    return g;
}

static void transfer_box_g(const orc_plan *p, const fan_grid *gr, int passive, int start[3], int maxI[3],
                           const int doseDims[3]) {
    const int sx = gr->sx, sy = gr->sy;
    /* rayIdx -> doseIdx in float, as the host does (Float3FromFanTransform::transformPoint) */
    const float xv[2] = {-(float)ORC_MAX_RAD, (float)(sx + ORC_MAX_RAD - 1)};
    const float yv[2] = {-(float)ORC_MAX_RAD, (float)(sy + ORC_MAX_RAD - 1)};
    const float zv[2] = {(float)p->beamFirstInside, (float)(passive - 1)};
    m3 M = m3_from9(p->g2dose);
    f3 off = mk3(p->g2doseOff[0], p->g2doseOff[1], p->g2doseOff[2]);
    f3 mx = mk3(-1.f, -1.f, -1.f), mn = mk3(100000.f, 100000.f, 100000.f);
    for (int a = 0; a < 2; ++a)
        for (int bb = 0; bb < 2; ++bb)
            for (int c = 0; c < 2; ++c) {
                f3 f = mk3(xv[c] * gr->fd[0] + gr->fc[0], yv[bb] * gr->fd[1] + gr->fc[1],
                           zv[a] * p->fanDelta[2] + p->fanCorner[2]);
                f.x *= (1.0f - f.z / p->sourceDist[0]);
                f.y *= (1.0f - f.z / p->sourceDist[1]);
                f3 q = add3(mv3(M, f), off);
                mx.x = fmaxf(mx.x, q.x); mx.y = fmaxf(mx.y, q.y); mx.z = fmaxf(mx.z, q.z);
                mn.x = fminf(mn.x, q.x); mn.y = fminf(mn.y, q.y); mn.z = fminf(mn.z, q.z);
            }
    int s0 = ((int)floorf(mn.x)) / 32 * 32;
    start[0] = s0 > 0 ? s0 : 0;
    start[1] = (int)floorf(mn.y) > 0 ? (int)floorf(mn.y) : 0;
    start[2] = (int)floorf(mn.z) > 0 ? (int)floorf(mn.z) : 0;
    maxI[0] = (int)ceilf(mx.x) < doseDims[0] - 1 ? (int)ceilf(mx.x) : doseDims[0] - 1;
    maxI[1] = (int)ceilf(mx.y) < doseDims[1] - 1 ? (int)ceilf(mx.y) : doseDims[1] - 1;
    maxI[2] = (int)ceilf(mx.z) < doseDims[2] - 1 ? (int)ceilf(mx.z) : doseDims[2] - 1;
}

static void transfer_box(const orc_plan *p, int passive, int start[3], int maxI[3], const int doseDims[3]) {
    fan_grid g = primary_grid(p);
    transfer_box_g(p, &g, passive, start, maxI, doseDims);
}

/* One layer's BEV of grid gr -> dose (primTransfDiv / nucTransfDiv, prim_transf_kernel.cu:11-78). */
static void transfer_g(orc_plan *p, const fan_grid *gr, int passive, const double *bev, double *dose,
                       const int doseDims[3]) {
    if (passive <= p->beamFirstInside) return;
    int st[3], mxI[3];
    transfer_box_g(p, gr, passive, st, mxI, doseDims);
    if (!(st[0] <= mxI[0] && st[1] <= mxI[1] && st[2] <= mxI[2])) return;
    /* the launch covers whole 32x8 blocks from the box start; threads stop only at the grid edge */
    int xEnd = st[0] + ((mxI[0] - st[0] + 1 + 31) / 32) * 32;
    int yEnd = st[1] + ((mxI[1] - st[1] + 1 + 7) / 8) * 8;
    if (xEnd > doseDims[0]) xEnd = doseDims[0];
    if (yEnd > doseDims[1]) yEnd = doseDims[1];
    /* doseIdx -> gantry: inverse of gantryToDoseIdx in float (Float3AffineTransform::inverse) */
    aff g2d;
    g2d.m = m3_from9(p->g2dose);
    g2d.v = mk3(p->g2doseOff[0], p->g2doseOff[1], p->g2doseOff[2]);
    aff d2g = aff_inv(g2d);
    double M[9], O[3];
    float t9[9];
    m3_to9(d2g.m, t9);
    for (int i = 0; i < 9; ++i) M[i] = t9[i];
    O[0] = d2g.v.x; O[1] = d2g.v.y; O[2] = d2g.v.z;
    const int bevDims[3] = {gr->sx + 2 * ORC_MAX_RAD, gr->sy + 2 * ORC_MAX_RAD, p->guaranteedPassive};
    const double shift[3] = {ORC_MAX_RAD, ORC_MAX_RAD, -(double)p->beamFirstInside};
#pragma omp parallel for schedule(static)
    for (int y = st[1]; y < yEnd; ++y)
        for (int x = st[0]; x < xEnd; ++x)
            for (int z = st[2]; z <= mxI[2]; ++z) {
                double g[3];
                for (int r = 0; r < 3; ++r) g[r] = M[3 * r] * x + M[3 * r + 1] * y + M[3 * r + 2] * z + O[r];
                double fx = g[0] / (1.0 - g[2] / p->sourceDist[0]);
                double fy = g[1] / (1.0 - g[2] / p->sourceDist[1]);
                double f[3] = {(fx - gr->fc[0]) / gr->fd[0] + shift[0],
                               (fy - gr->fc[1]) / gr->fd[1] + shift[1],
                               (g[2] - p->fanCorner[2]) / p->fanDelta[2] + shift[2]};
                /* prim_transf_kernel.cu:28-40: pos = fanIdx + 0.5; skip negatives; add if > 0 */
                if (f[0] + 0.5 < 0.0 || f[1] + 0.5 < 0.0 || f[2] + 0.5 < 0.0) continue;
                double v = sample3d_border_d(bev, bevDims, f[0], f[1], f[2]);
                if (v > 0.0) dose[((size_t)z * doseDims[1] + y) * doseDims[0] + x] += v;
            }
}

void orc_transfer(orc_plan *p, int l, const double *bev, double *dose, const int doseDims[3]) {
    fan_grid g = primary_grid(p);
    transfer_g(p, &g, p->layerPassive[l], bev, dose, doseDims);
}

/* ------------------------------------------------------------------------------------
 * whole beam
 * ---------------------------------------------------------------------------------- */
void orc_trace_entry_margin(const orc_plan *p, const float *ct, const int ctDims[3], double *margin) {
    const int rx = p->rayDims[0], ry = p->rayDims[1], K = p->steps;
    if (p->weqVolume) {
        for (size_t r = 0; r < (size_t)rx * ry; ++r) margin[r] = INFINITY;
        return;
    }
    fan_geom g;
    fan_geom_from(p->g2im, p->g2imOff, p, &g);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ry; ++j) {
        for (int i = 0; i < rx; ++i) {
            double a[3], bpt[3];
            fan_to_index(&g, i, j, 0, a);
            fan_to_index(&g, i, j, 1, bpt);
            double inc[3] = {bpt[0] - a[0], bpt[1] - a[1], bpt[2] - a[2]};
            /* the cumulative sum never decreases: the decision is fixed by the last step below the
             * threshold and the first at or above it */
            double c = 0.0, below = -INFINITY, m = INFINITY;
            for (int k = 0; k < K; ++k) {
                c += orc_sample3d_border(ct, ctDims, a[0] + k * inc[0], a[1] + k * inc[1], a[2] + k * inc[2]);
                if (c < 150.0) { below = c; continue; }
                m = c - 150.0;
                break;
            }
            if (below > -INFINITY && 150.0 - below < m) m = 150.0 - below;
            margin[(size_t)j * rx + i] = m;
        }
    }
}

int orc_beam_dose(const float *ct, const int ctDims[3], const float ctRes[3], const float ctCorner[3],
                  double *dose, const int doseDims[3], const float doseRes[3], const float doseCorner[3],
                  const orc_beam *b, const orc_energy *en, orc_plan *planOut) {
    return orc_beam_dose_entry(ct, ctDims, ctRes, ctCorner, dose, doseDims, doseRes, doseCorner, b, en, NULL,
                               planOut);
}

int orc_beam_dose_entry(const float *ct, const int ctDims[3], const float ctRes[3], const float ctCorner[3],
                        double *dose, const int doseDims[3], const float doseRes[3], const float doseCorner[3],
                        const orc_beam *b, const orc_energy *en, const int *entryOverride, orc_plan *planOut) {
    orc_plan *p = planOut ? planOut : (orc_plan *)malloc(sizeof(orc_plan));
    if (orc_plan_beam(ctDims, ctRes, ctCorner, doseDims, doseRes, doseCorner, b, en, p) != 0) {
        if (!planOut) free(p);
        return -1;
    }
    const int rx = p->rayDims[0], ry = p->rayDims[1], K = p->steps, L = b->numLayers;
    const int sx = p->superpDims[0], sy = p->superpDims[1];
    const size_t rplane = (size_t)rx * ry, splane = (size_t)sx * sy;
    double *density = (double *)malloc(sizeof(double) * rplane * K);
    double *wepl = (double *)malloc(sizeof(double) * rplane * K);
    int *fi = (int *)malloc(sizeof(int) * rplane), *fo = (int *)malloc(sizeof(int) * rplane);
    double *weplMin = (double *)malloc(sizeof(double) * K);
    orc_trace(p, ct, ctDims, en, density, wepl, fi, fo);
    if (entryOverride)
        for (size_t r = 0; r < rplane; ++r)
            if (entryOverride[r] >= 0) fi[r] = entryOverride[r];
    orc_beam_cutoffs(p, b, en, wepl, fi, fo, weplMin);
    double *rw = (double *)calloc(rplane * L, sizeof(double));
    if (!p->legacyPath) {
        orc_conv_lattice(p, b, en, rw);
    } else {
        const int cd[3] = {rx, ry, L};
        const float bdG[3] = {0.f, 0.f, -1.f}, bxG[3] = {1.f, 0.f, 0.f}, byG[3] = {0.f, -1.f, 0.f};
        float sad = (b->sad > 0.0f) ? b->sad : b->sourceDist[0];
        const float srcG[3] = {0.f, 0.f, sad};
        double *cpb = (double *)malloc(sizeof(double) * rplane * L);
        orc_subspot_to_cpb(b->subspotData, L, b->maxSubspots, p->cpbCorner, p->cpbRes, cd, bdG, bxG, byG, srcG, sad, cpb);
        for (int l = 0; l < L; ++l)
            orc_cpb_to_rays(cpb, cd, p->cpbCorner, p->cpbRes, rw + (size_t)l * rplane, p->rayDims, p->cpbCorner, p->cpbRes, l);
        free(cpb);
    }
    orc_layer_params(p, b, en, weplMin);
    double *idd = (double *)malloc(sizeof(double) * rplane * K);
    double *rs = (double *)malloc(sizeof(double) * rplane * K);
    double *iddS = NULL, *rsS = NULL;
    int *fp = (int *)malloc(sizeof(int) * rplane);
    const size_t bevN = (size_t)(sx + 64) * (sy + 64) * (p->guaranteedPassive > 0 ? p->guaranteedPassive : 1);
    double *bev = (double *)malloc(sizeof(double) * bevN);
    if (p->legacyPath) {
        iddS = (double *)malloc(sizeof(double) * splane * K);
        rsS = (double *)malloc(sizeof(double) * splane * K);
    }
    /* halo branch (wrapper:3413-3450, 4062-4470, 5102-5215, 5510-5600) */
    int *rayToNuc = NULL;
    double *nucW = NULL, *nucIdd = NULL, *nucRs = NULL, *nucBev = NULL;
    size_t nucPlane = 0, nucBevN = 0;
    int haloFail = 0;
    if (b->nuclear) {
        if (p->legacyPath || !en->nucWeight || !en->nucSqSigma) {
            snprintf(p->error, sizeof(p->error), p->legacyPath
                         ? "nuclear_correction=true requires raw spot lattice inputs; legacy CPB fallback cannot drive halo mode"
                         : "nuclear_correction=true requested, but the energy struct does not carry nuclear tables");
            haloFail = 1;
        } else {
            rayToNuc = (int *)malloc(sizeof(int) * rplane);
            nucW = (double *)malloc(sizeof(double) * (size_t)L * (((size_t)p->spotDims[0] + 31) / 32 * 32) *
                                    (((size_t)p->spotDims[1] + 7) / 8 * 8));
            if (orc_halo_plan(p, b, rayToNuc, nucW) != 0) haloFail = 1;
            nucPlane = (size_t)p->nucDims[0] * p->nucDims[1];
            nucIdd = (double *)malloc(sizeof(double) * nucPlane * K);
            nucRs = (double *)malloc(sizeof(double) * nucPlane * K);
            nucBevN = (size_t)(p->nucDims[0] + 64) * (p->nucDims[1] + 64) * (p->guaranteedPassive > 0 ? p->guaranteedPassive : 1);
            nucBev = (double *)malloc(sizeof(double) * nucBevN);
        }
    }
    fan_grid nucGrid = {p->nucDims[0], p->nucDims[1], {p->nucOffset[0], p->nucOffset[1]},
                        {p->nucDelta[0], p->nucDelta[1]}};
    for (int l = 0; l < L && !haloFail; ++l) {
        if (b->nuclear) {
            for (size_t i = 0; i < nucPlane * K; ++i) { nucIdd[i] = 0.0; nucRs[i] = INFINITY; }
            orc_fill_halo(p, l, en, density, wepl, rw + (size_t)l * rplane, fi, fo, idd, rs, fp, b->nuclear, rayToNuc,
                          nucW + (size_t)l * nucPlane, nucIdd, nucRs);
        } else {
            orc_fill(p, l, en, density, wepl, rw + (size_t)l * rplane, fi, fo, idd, rs, fp);
        }
        /* the CarbonPBS profile hooks rewrite the primary field (wrapper:4604-4832) */
        orc_profile_hooks(p, l, wepl, idd, rs);
        int mxp = INT32_MIN;
        for (size_t r = 0; r < rplane; ++r) if (fp[r] > mxp) mxp = fp[r];
        /* wrapper:4984-4996 */
        int pas = mxp;
        if (pas < p->beamFirstInside + 1) pas = p->beamFirstInside + 1;
        if (pas < 0) pas = p->beamFirstInside + 1;
        if (pas > p->afterLast[l]) pas = p->afterLast[l];
        if (pas > K) pas = K;
        p->layerPassive[l] = pas;
        const double *iddUse = idd, *rsUse = rs;
        if (p->legacyPath) {
            /* padRayIddSigmaKernel (wrapper:2027-2053) over the steps the superposition reads */
            for (int k = p->beamFirstInside; k < pas; ++k)
                for (int y = 0; y < sy; ++y)
                    for (int x = 0; x < sx; ++x) {
                        size_t d = (size_t)k * splane + (size_t)y * sx + x;
                        if (x < rx && y < ry) {
                            size_t s = (size_t)k * rplane + (size_t)y * rx + x;
                            iddS[d] = idd[s]; rsS[d] = rs[s];
                        } else { iddS[d] = 0.0; rsS[d] = INFINITY; }
                    }
            iddUse = iddS; rsUse = rsS;
        }
        memset(bev, 0, sizeof(double) * bevN);
        orc_superpose(p, l, iddUse, rsUse, bev);
        orc_transfer(p, l, bev, dose, doseDims);
        if (b->nuclear) {
            int batch[ORC_MAX_RAD + 1];
            memset(nucBev, 0, sizeof(double) * nucBevN);
            if (superpose_grid(p->beamFirstInside, pas, p->nucDims[0], p->nucDims[1], nucIdd, nucRs, nucBev,
                               p->nucRadHist[l], batch, 1) != 0) {
                snprintf(p->error, sizeof(p->error), "nuclear superposition kernel radius exceeded %d (layer %d)",
                         ORC_MAX_RAD, l);
                haloFail = 1;
                break;
            }
            transfer_g(p, &nucGrid, pas, nucBev, dose, doseDims);
        }
    }
    free(rayToNuc); free(nucW); free(nucIdd); free(nucRs); free(nucBev);
    /* union transfer box over layers, reported for tests */
    {
        int maxPas = p->beamFirstInside;
        for (int l = 0; l < L; ++l) if (p->layerPassive[l] > maxPas) maxPas = p->layerPassive[l];
        if (maxPas > p->beamFirstInside) transfer_box(p, maxPas, p->transferStart, p->transferMax, doseDims);
    }
    free(density); free(wepl); free(fi); free(fo); free(weplMin); free(rw);
    free(idd); free(rs); free(fp); free(bev); free(iddS); free(rsS);
    if (!planOut) free(p);
    return haloFail ? -1 : 0;
}

/* ------------------------------------------------------------------------------------
 * adjoint with respect to the spot weights (lattice path, halo off)
 *
 * For fixed live masks, tile radii and step ranges (all of which depend on the CT, the
 * sigmas and the sign of the weights, not on their size) the dose of orc_beam_dose is linear
 * in the spot weights: dose = T S F C w (C: spot lattice -> ray weights, F: IDD = ray weight
 * x unit IDD, S: superposition, T: BEV -> dose transfer). The adjoint applies the transposes
 * in reverse order to dJ/dDose. Each transpose walks the same index pairs as its forward
 * loop and scatters, so the pair sets are identical.
 * ---------------------------------------------------------------------------------- */
/* transpose of conv1d_gather: gIn[ci] += tap(o, ci) gOut[o] */
static void conv1d_gather_adj(const double *gOut, int outW, int outStride, double *gIn, int inW, int inStride,
                              double sigma, double pixelSp, double inOutOffset, double inOutDelta) {
    const double sEff = sigma / pixelSp;
    const double rS = 1.0 / (sqrt(2.0) * sEff);
    const double reach = 3.0 * sEff + 0.5;
    for (int o = 0; o < outW; ++o) {
        const double go = gOut[(size_t)o * outStride];
        int ci = (int)ceil(((double)o - reach - inOutOffset) / inOutDelta);
        double dist = ci * inOutDelta + inOutOffset - (double)o;
        while (dist < reach) {
            if (ci >= 0 && ci < inW)
                gIn[(size_t)ci * inStride] += 0.5 * (erf((dist + 0.5) * rS) - erf((dist - 0.5) * rS)) * go;
            ++ci;
            dist = ci * inOutDelta + inOutOffset - (double)o;
        }
    }
}

/* transpose of conv_plane: gSpot [sy][sx] += C^T (scale gRays) */
static void conv_plane_adj(const double *gRays, int rx, int ry, double *gSpot, int sx, int sy, double sigX,
                           double sigY, const orc_plan *p, float offX, float offY, double pxMultX, double pxMultY,
                           double scale) {
    double *gTmp = (double *)malloc(sizeof(double) * (size_t)rx * ry);
    double *gInterm = (double *)calloc((size_t)rx * sy, sizeof(double));
    float dX = p->spotDelta[0] / p->cpbRes[0], dY = p->spotDelta[1] / p->cpbRes[1];
    float oX = (offX - p->cpbCorner[0]) / p->cpbRes[0], oY = (offY - p->cpbCorner[1]) / p->cpbRes[1];
    for (size_t i = 0; i < (size_t)rx * ry; ++i) gTmp[i] = scale * gRays[i];
    for (int x = 0; x < rx; ++x)
        conv1d_gather_adj(gTmp + x, ry, rx, gInterm + x, sy, rx, sigY, (double)p->cpbRes[1] * pxMultY, oY, dY);
    for (int y = 0; y < sy; ++y)
        conv1d_gather_adj(gInterm + (size_t)y * rx, rx, 1, gSpot + (size_t)y * sx, sx, 1, sigX,
                          (double)p->cpbRes[0] * pxMultX, oX, dX);
    free(gTmp); free(gInterm);
}

/* transpose of orc_conv_lattice for one layer: gSpot [sy][sx] from gRays [ry][rx] */
static void conv_lattice_adj(const orc_plan *p, const orc_beam *b, const orc_energy *en, int l, const double *gRays,
                             double *gSpot) {
    const int rx = p->rayDims[0], ry = p->rayDims[1], sx = p->spotDims[0], sy = p->spotDims[1];
    const float entryZ = (float)p->beamFirstInside * p->fanDelta[2] + p->fanCorner[2];
    const double pxX = 1.0f - entryZ / p->sourceDist[0], pxY = 1.0f - entryZ / p->sourceDist[1];
    float eIdx = orc_find_decimal_ordered(en->energies, en->nEnergies, b->energies[l]);
    float peak = orc_vector_interpolate(en->peaks, en->nEnergies, eIdx) * p->energyDepthToMm;
    float ca, cb;
    air_coefs(peak, &ca, &cb);
    int done = 0;
    if (b->useSubspotConvolution && b->maxSubspots > 0) {
        for (int s = 0; s < b->maxSubspots; ++s) {
            const float *ss = b->subspotData + ((size_t)l * b->maxSubspots + s) * 5;
            if (!(ss[2] > 1.0e-8f)) continue;
            float sgx = sigma_to_mm(p->lenToMm, ss[3]), sgy = sigma_to_mm(p->lenToMm, ss[4]);
            if (!(sgx > 1.0e-8f) || !(sgy > 1.0e-8f)) continue;
            float vx = ca * entryZ * entryZ + cb * entryZ + sgx * sgx;
            float vy = ca * entryZ * entryZ + cb * entryZ + sgy * sgy;
            double ex = sqrt((double)(vx > 0.f ? vx : 0.f)), ey = sqrt((double)(vy > 0.f ? vy : 0.f));
            conv_plane_adj(gRays, rx, ry, gSpot, sx, sy, ex, ey, p, p->spotOffset[0] + ss[0] / p->lenToMm,
                           p->spotOffset[1] + ss[1] / p->lenToMm, pxX, pxY, (double)ss[2]);
            done = 1;
        }
    }
    if (!done) {
        float sgx = sigma_to_mm(p->lenToMm, b->spotSigmas[2 * l]);
        float sgy = sigma_to_mm(p->lenToMm, b->spotSigmas[2 * l + 1]);
        float vx = ca * entryZ * entryZ + cb * entryZ + sgx * sgx;
        float vy = ca * entryZ * entryZ + cb * entryZ + sgy * sgy;
        double ex = sqrt((double)(vx > 0.f ? vx : 0.f)), ey = sqrt((double)(vy > 0.f ? vy : 0.f));
        conv_plane_adj(gRays, rx, ry, gSpot, sx, sy, ex, ey, p, p->spotOffset[0], p->spotOffset[1], pxX, pxY, 1.0);
    }
}

/* transpose of sample3d_border_d: scatter g into the in-range corners */
static void sample3d_border_adj(double *vol, const int dims[3], double x, double y, double z, double g) {
    double fx = floor(x), fy = floor(y), fz = floor(z);
    long x0 = (long)fx, y0 = (long)fy, z0 = (long)fz;
    double tx = x - fx, ty = y - fy, tz = z - fz;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                long xi = x0 + dx, yi = y0 + dy, zi = z0 + dz;
                if (xi < 0 || yi < 0 || zi < 0 || xi >= dims[0] || yi >= dims[1] || zi >= dims[2]) continue;
                const double w = (dx ? tx : 1.0 - tx) * (dy ? ty : 1.0 - ty) * (dz ? tz : 1.0 - tz);
                vol[((size_t)zi * dims[1] + yi) * dims[0] + xi] += w * g;
            }
}

/* transpose of transfer_g: gBev += T^T gDose (the forward adds the sample when it is positive;
 * samples of a non-negative field are non-negative, so the map is linear) */
static void transfer_g_adj(orc_plan *p, const fan_grid *gr, int passive, const double *gDose, double *gBev,
                           const int doseDims[3]) {
    if (passive <= p->beamFirstInside) return;
    int st[3], mxI[3];
    transfer_box_g(p, gr, passive, st, mxI, doseDims);
    if (!(st[0] <= mxI[0] && st[1] <= mxI[1] && st[2] <= mxI[2])) return;
    int xEnd = st[0] + ((mxI[0] - st[0] + 1 + 31) / 32) * 32;
    int yEnd = st[1] + ((mxI[1] - st[1] + 1 + 7) / 8) * 8;
    if (xEnd > doseDims[0]) xEnd = doseDims[0];
    if (yEnd > doseDims[1]) yEnd = doseDims[1];
    aff g2d;
    g2d.m = m3_from9(p->g2dose);
    g2d.v = mk3(p->g2doseOff[0], p->g2doseOff[1], p->g2doseOff[2]);
    aff d2g = aff_inv(g2d);
    double M[9], O[3];
    float t9[9];
    m3_to9(d2g.m, t9);
    for (int i = 0; i < 9; ++i) M[i] = t9[i];
    O[0] = d2g.v.x; O[1] = d2g.v.y; O[2] = d2g.v.z;
    const int bevDims[3] = {gr->sx + 2 * ORC_MAX_RAD, gr->sy + 2 * ORC_MAX_RAD, p->guaranteedPassive};
    const double shift[3] = {ORC_MAX_RAD, ORC_MAX_RAD, -(double)p->beamFirstInside};
    for (int y = st[1]; y < yEnd; ++y)
        for (int x = st[0]; x < xEnd; ++x)
            for (int z = st[2]; z <= mxI[2]; ++z) {
                const double gv = gDose[((size_t)z * doseDims[1] + y) * doseDims[0] + x];
                if (gv == 0.0) continue;
                double g[3];
                for (int r = 0; r < 3; ++r) g[r] = M[3 * r] * x + M[3 * r + 1] * y + M[3 * r + 2] * z + O[r];
                double fx = g[0] / (1.0 - g[2] / p->sourceDist[0]);
                double fy = g[1] / (1.0 - g[2] / p->sourceDist[1]);
                double f[3] = {(fx - gr->fc[0]) / gr->fd[0] + shift[0], (fy - gr->fc[1]) / gr->fd[1] + shift[1],
                               (g[2] - p->fanCorner[2]) / p->fanDelta[2] + shift[2]};
                if (f[0] + 0.5 < 0.0 || f[1] + 0.5 < 0.0 || f[2] + 0.5 < 0.0) continue;
                sample3d_border_adj(gBev, bevDims, f[0], f[1], f[2], gv);
            }
}

/* transpose of superpose_grid (same radii and batching): gIdd[k][pix] = sum over the source's
 * footprint of e_|j| e_|i| gBev, for the pixels with a non-zero IDD (the only ones the chain
 * to the weights needs) */
static void superpose_grid_adj(int k0, int k1, int sx, int sy, const double *idd, const double *rsigma,
                               const double *gBev, double *gIdd) {
    const int ntx = sx / ORC_TILE_X, nty = sy / ORC_TILE_Y;
    const int bx = sx + 2 * ORC_MAX_RAD, by = sy + 2 * ORC_MAX_RAD;
    (void)by;
    if (k1 <= k0) return;
    int hist[ORC_MAX_RAD + 2], batch[ORC_MAX_RAD + 1], effRad[ORC_MAX_RAD + 2];
    memset(hist, 0, sizeof(hist));
    const int nk = k1 - k0;
    int *rad = (int *)malloc(sizeof(int) * (size_t)nk * ntx * nty);
    for (int k = k0; k < k1; ++k)
        for (int ty = 0; ty < nty; ++ty)
            for (int tx = 0; tx < ntx; ++tx) {
                double m = NAN;
                for (int y = 0; y < ORC_TILE_Y; ++y)
                    for (int x = 0; x < ORC_TILE_X; ++x)
                        m = fmin(m, rsigma[((size_t)k * sy + ty * ORC_TILE_Y + y) * sx + tx * ORC_TILE_X + x]);
                int r = orc_tile_radius_f32((float)m);
                rad[((size_t)(k - k0) * nty + ty) * ntx + tx] = r;
                hist[r]++;
            }
    orc_batch_bins(hist, batch, effRad);
    double e[ORC_MAX_RAD + 1];
    for (int k = k0; k < k1; ++k) {
        const double *plane = gBev + (size_t)k * bx * (sy + 2 * ORC_MAX_RAD);
        for (int ty = 0; ty < nty; ++ty)
            for (int tx = 0; tx < ntx; ++tx) {
                const int R = effRad[rad[((size_t)(k - k0) * nty + ty) * ntx + tx]];
                for (int y = 0; y < ORC_TILE_Y; ++y)
                    for (int x = 0; x < ORC_TILE_X; ++x) {
                        const size_t ii = ((size_t)k * sy + ty * ORC_TILE_Y + y) * sx + tx * ORC_TILE_X + x;
                        if (idd[ii] == 0.0) { gIdd[ii] = 0.0; continue; }
                        const double r = rsigma[ii];
                        double en = erf(r * 0.5), eo = -en;
                        for (int i = 0; i <= R; ++i) {
                            e[i] = 0.5 * (en - eo);
                            eo = en;
                            en = erf(r * ((double)i + 1.5));
                        }
                        const int X = tx * ORC_TILE_X + x + ORC_MAX_RAD;
                        const int Y = ty * ORC_TILE_Y + y + ORC_MAX_RAD;
                        double acc = 0.0;
                        for (int j = -R; j <= R; ++j) {
                            const double *row = plane + (size_t)(Y + j) * bx + X;
                            double rsum = 0.0;
                            for (int i = -R; i <= R; ++i) rsum += e[i < 0 ? -i : i] * row[i];
                            acc += e[j < 0 ? -j : j] * rsum;
                        }
                        gIdd[ii] = acc;
                    }
            }
    }
    free(rad);
}

int orc_beam_adjoint(const float *ct, const int ctDims[3], const float ctRes[3], const float ctCorner[3],
                     const double *gDose, const int doseDims[3], const float doseRes[3], const float doseCorner[3],
                     const orc_beam *b, const orc_energy *en, double *gSpots, orc_plan *planOut) {
    orc_plan *p = planOut ? planOut : (orc_plan *)malloc(sizeof(orc_plan));
    if (orc_plan_beam(ctDims, ctRes, ctCorner, doseDims, doseRes, doseCorner, b, en, p) != 0) {
        if (!planOut) free(p);
        return -1;
    }
    if (p->legacyPath || b->nuclear) {
        snprintf(p->error, sizeof(p->error), "the adjoint covers the spot-lattice path without the halo branch");
        if (!planOut) free(p);
        return -1;
    }
    const int rx = p->rayDims[0], ry = p->rayDims[1], K = p->steps, L = b->numLayers;
    const int sx = p->superpDims[0], sy = p->superpDims[1];
    const int spx = p->spotDims[0], spy = p->spotDims[1];
    const size_t rplane = (size_t)rx * ry;
    double *density = (double *)malloc(sizeof(double) * rplane * K);
    double *wepl = (double *)malloc(sizeof(double) * rplane * K);
    int *fi = (int *)malloc(sizeof(int) * rplane), *fo = (int *)malloc(sizeof(int) * rplane);
    double *weplMin = (double *)malloc(sizeof(double) * K);
    orc_trace(p, ct, ctDims, en, density, wepl, fi, fo);
    orc_beam_cutoffs(p, b, en, wepl, fi, fo, weplMin);
    double *rw = (double *)calloc(rplane * L, sizeof(double));
    orc_conv_lattice(p, b, en, rw);
    orc_layer_params(p, b, en, weplMin);
    double *idd = (double *)malloc(sizeof(double) * rplane * K);
    double *rs = (double *)malloc(sizeof(double) * rplane * K);
    double *gIdd = (double *)malloc(sizeof(double) * rplane * K);
    double *gW = (double *)malloc(sizeof(double) * rplane);
    double *gSpot = (double *)calloc((size_t)spx * spy * L, sizeof(double));
    int *fp = (int *)malloc(sizeof(int) * rplane);
    const size_t bevN = (size_t)(sx + 64) * (sy + 64) * (p->guaranteedPassive > 0 ? p->guaranteedPassive : 1);
    double *gBev = (double *)malloc(sizeof(double) * bevN);
    fan_grid gr = primary_grid(p);
    for (int l = 0; l < L; ++l) {
        orc_fill(p, l, en, density, wepl, rw + (size_t)l * rplane, fi, fo, idd, rs, fp);
        int mxp = INT32_MIN;
        for (size_t r = 0; r < rplane; ++r) if (fp[r] > mxp) mxp = fp[r];
        int pas = mxp;
        if (pas < p->beamFirstInside + 1) pas = p->beamFirstInside + 1;
        if (pas < 0) pas = p->beamFirstInside + 1;
        if (pas > p->afterLast[l]) pas = p->afterLast[l];
        if (pas > K) pas = K;
        p->layerPassive[l] = pas;
        memset(gBev, 0, sizeof(double) * bevN);
        transfer_g_adj(p, &gr, pas, gDose, gBev, doseDims);
        superpose_grid_adj(p->beamFirstInside, pas, sx, sy, idd, rs, gBev, gIdd);
        /* IDD = W x (unit IDD) on live rays: dJ/dW = sum_k dJ/dIDD IDD / W */
        const double *W = rw + (size_t)l * rplane;
        for (size_t r = 0; r < rplane; ++r) {
            double acc = 0.0;
            if (W[r] >= 1e-6)
                for (int k = p->beamFirstInside; k < pas; ++k) {
                    const size_t ii = (size_t)k * rplane + r;
                    if (idd[ii] != 0.0) acc += gIdd[ii] * idd[ii];
                }
            gW[r] = W[r] >= 1e-6 ? acc / W[r] : 0.0;
        }
        conv_lattice_adj(p, b, en, l, gW, gSpot + (size_t)l * spx * spy);
    }
    /* spots -> their lattice cells (build_lattice's binning) */
    int off = 0;
    for (int l = 0; l < L; ++l) {
        for (int i = 0; i < b->layerSpotCounts[l]; ++i) {
            const int s = off + i;
            const float x = b->spotPositions[2 * s], y = b->spotPositions[2 * s + 1];
            const float fx = p->spotDelta[0] != 0.0f ? (x - p->spotOffset[0]) / p->spotDelta[0] : 0.0f;
            const float fy = p->spotDelta[1] != 0.0f ? (y - p->spotOffset[1]) / p->spotDelta[1] : 0.0f;
            const int ix = (int)lroundf(fx), iy = (int)lroundf(fy);
            gSpots[s] = (ix >= 0 && ix < spx && iy >= 0 && iy < spy) ? gSpot[((size_t)l * spy + iy) * spx + ix] : 0.0;
        }
        off += b->layerSpotCounts[l];
    }
    free(density); free(wepl); free(fi); free(fo); free(weplMin); free(rw);
    free(idd); free(rs); free(gIdd); free(gW); free(gSpot); free(fp); free(gBev);
    if (!planOut) free(p);
    return 0;
}
