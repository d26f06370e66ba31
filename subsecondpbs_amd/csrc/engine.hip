// Engine orchestration and C ABI of the MI355X pencil-beam final-dose path.
//
// One beam is a fixed sequence of launches on the engine's stream with no host round trip:
// trace -> trace reductions -> cut-offs -> lattice convolution (or subspot->CPB) -> IDD/sigma
// fill -> layer finalisation -> tile radii / bins / items -> superposition (8 radius
// classes) -> block merge -> BEV->dose transfer. Device scalars (entry step, cut-offs,
// passive steps, transfer box) are produced and consumed on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pbs_abi.h"
#include "device.h"
#include "kernels.h"
#include "plan.h"

namespace {

thread_local std::string g_lastError;

int set_error(int code, const std::string &msg) {
    g_lastError = msg;
    return code;
}

#define HIP_TRY(expr)                                                                                      \
    do {                                                                                                   \
        hipError_t _e = (expr);                                                                            \
        if (_e != hipSuccess)                                                                              \
            return set_error(PBS_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e));            \
    } while (0)

// Bumped whenever any device buffer of the library is freed or moved: a captured HIP graph
// holds raw addresses and is replayed only while this count is the one it was captured at.
std::atomic<unsigned long long> g_bufGen{0};

// Bump allocator over a device scratch region, active on the calling thread for the duration of
// a no-allocation entry point: DevBuf::ensure then carves from it instead of calling hipMalloc.
struct Arena {
    unsigned char *base = nullptr;  // device address of the scratch
    size_t cap = 0, used = 0;
    bool overflow = false;
};
thread_local Arena *t_arena = nullptr;

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    bool borrowed = false;  // p lies in a no-allocation call's scratch (never freed)
    T *parked = nullptr;    // a too-small own allocation replaced during such a call, freed later
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p && !borrowed) {
            (void)hipFree(p);
            g_bufGen.fetch_add(1);
        }
        if (parked) {
            (void)hipFree(parked);
            g_bufGen.fetch_add(1);
        }
        p = nullptr;
        parked = nullptr;
        borrowed = false;
        n = 0;
    }
    hipError_t ensure(size_t count) {
        count = std::max<size_t>(count, 1);
        if (Arena *a = t_arena) {
            // an own allocation that is large enough is used as it is; anything else is carved
            // from the scratch (a carve of an earlier call is stale: the scratch is re-carved on
            // every call)
            if (p && !borrowed && count <= n) return hipSuccess;
            const size_t bytes = (count * sizeof(T) + 255) & ~size_t(255);
            if (a->used + bytes > a->cap) {
                a->overflow = true;
                return hipErrorOutOfMemory;
            }
            if (p && !borrowed) {
                if (parked) return hipErrorOutOfMemory;  // cannot happen: a parked buffer's p is borrowed
                parked = p;
            }
            p = reinterpret_cast<T *>(a->base + a->used);
            a->used += bytes;
            n = count;
            borrowed = true;
            g_bufGen.fetch_add(1);
            return hipSuccess;
        }
        if (count <= n && p && !borrowed) return hipSuccess;
        release();
        hipError_t e = hipMalloc(reinterpret_cast<void **>(&p), count * sizeof(T));
        if (e == hipSuccess) n = count;
        return e;
    }
};

}  // namespace

using namespace pbs;

// Device scratch of pbs_engine_beam_dose_noalloc: every per-beam buffer of such a call is carved
// from it, so the call allocates nothing. Each device holds its own copy from the library's load
// until the process ends. Host code never uses the array's name, only the device address that
// hipGetSymbolAddress returns for the current device (g_scratchDev).
constexpr size_t kNoAllocScratchBytes = size_t(256) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratchMu;
static unsigned char *g_scratchDev[kMaxDevices] = {};
static bool g_scratchTried[kMaxDevices] = {};

// The scratch's device address on `dev`, looked up once per device with that device current
// (engine creation); nullptr when the lookup failed.
static unsigned char *scratch_on_device(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratchMu);
    if (!g_scratchTried[dev]) {
        g_scratchTried[dev] = true;
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) == hipSuccess)
            g_scratchDev[dev] = static_cast<unsigned char *>(ptr);
        else
            (void)hipGetLastError();
    }
    return g_scratchDev[dev];
}

// Buffers of one superposition grid (tiles, work lists, item blocks, BEV planes, groups).
struct GridWork {
    DevBuf<int> tileRad, hist, effRad, itemR, classCount, tailPlaneGroup, tailPlaneK, tailItemR, planeRM, slotLayers;
    DevBuf<int2> classList;
    DevBuf<float> slots, tailSlots, bev, tailBev;
    DevBuf<GroupScalars> groupScal;
    int maxItems = 0, maxUnits = 0, tailCap = 0, nt = 0;

    hipError_t ensure(int L, int K, int sx, int sy) {
        nt = (sx / 32) * (sy / 8);
        maxItems = K * nt;
        // a group's tail is at most min(beamFirstInside + 1, K - beamFirstInside) <= (K + 1) / 2
        // steps long, and there are at most min(L, K) groups
        tailCap = std::min(L, K) * std::max(1, (K + 1) / 2);
        maxUnits = maxItems + tailCap * nt;
        const size_t bevPlane = size_t(sx + 64) * (sy + 64);
        hipError_t r;
        if ((r = tileRad.ensure(size_t(L) * K * nt)) != hipSuccess) return r;
        if ((r = slotLayers.ensure(size_t(L) * K * nt)) != hipSuccess) return r;
        if ((r = hist.ensure(size_t(L) * (kMaxRadDev + 2))) != hipSuccess) return r;
        if ((r = effRad.ensure(size_t(L) * (kMaxRadDev + 2))) != hipSuccess) return r;
        if ((r = itemR.ensure(maxItems)) != hipSuccess) return r;
        if ((r = classCount.ensure(16)) != hipSuccess) return r;
        if ((r = classList.ensure(size_t(kNumClasses) * maxUnits)) != hipSuccess) return r;
        if ((r = slots.ensure(size_t(maxItems) * kSlotFloats)) != hipSuccess) return r;
        if ((r = bev.ensure(size_t(K) * bevPlane)) != hipSuccess) return r;
        if ((r = groupScal.ensure(L)) != hipSuccess) return r;
        if ((r = tailPlaneGroup.ensure(tailCap)) != hipSuccess) return r;
        if ((r = tailPlaneK.ensure(tailCap)) != hipSuccess) return r;
        if ((r = tailItemR.ensure(size_t(tailCap) * nt)) != hipSuccess) return r;
        if ((r = planeRM.ensure(size_t(K) + tailCap)) != hipSuccess) return r;
        if ((r = tailSlots.ensure(size_t(tailCap) * nt * kSlotFloats)) != hipSuccess) return r;
        return tailBev.ensure(size_t(tailCap) * bevPlane);
    }
};

struct pbs_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    // LUTs (device) and host copies used by planning
    DevBuf<float> cidd, dens, sp, rrl, nucW, nucS2;
    std::vector<float> hEnergies, hPeaks, hScales, hDens, hSp;
    pbs_energy_struct energy{};
    PlanEnv env{};  // the environment's unit / input-type overrides, read when the engine is created
    LutDev lut{};
    // CT
    DevBuf<float> ct;
    const float *ctPtr = nullptr;
    int ctDims[3] = {0, 0, 0};
    float ctRes[3] = {0, 0, 0}, ctCorner[3] = {0, 0, 0};
    // dose grid
    int doseDims[3] = {0, 0, 0};
    float doseRes[3] = {0, 0, 0}, doseCorner[3] = {0, 0, 0};
    // per-beam work buffers
    DevBuf<float> density, wepl, weplMin, spot, interm, rayW, subspots, subEntry, cpb, traceScratch, weq;
    DevBuf<float2> iddSigma;
    DevBuf<int> fi, fo, firstPassive;
    // superposition -> merge -> transfer state of the primary (ray) grid and the halo grid
    GridWork prim, halo;
    // halo branch: spot-centre map, halo weights and fields, its own layer/beam scalars
    int nuclearModel = 0;  // 0 off, 1 GAUSS_FIT, 2 SOUKUP, 3 FLUKA
    HaloPlan haloPlan;
    DevBuf<int> rayToNuc, haloErr;
    DevBuf<float> haloW;
    DevBuf<float2> haloIddSigma;
    DevBuf<LayerScalars> haloLayerScal;
    DevBuf<BeamScalars> haloBeamScal;
    FanGeom haloGeom{};
    bool lastHalo = false, haloUsed = false;
    DevBuf<LayerHost> layerHost;
    DevBuf<LayerScalars> layerScal;
    DevBuf<BeamScalars> beamScal;
    // last beam (debug fetch)
    BeamPlan plan;
    FanGeom geom{};
    bool haveLast = false;
    // host staging of the last beam's uploads; the next beam waits for uploadDone before reuse
    std::vector<LayerHost> hostLayers;
    // pinned staging of a beam's uploads (per-layer values, subspot table, spot lattice, halo
    // arrays): copies from it are plain DMA, queued without holding the host
    char *pin = nullptr;
    char *pinDev = nullptr;  // the staging's device-visible address (read by k_upload)
    size_t pinBytes = 0;
    hipEvent_t uploadDone = nullptr;
    // radius classes of the superposition run concurrently on the engine stream and these
    // side streams (fork after the work lists, join before the merge)
    hipStream_t side[kSideStreams] = {};
    hipEvent_t forkEv = nullptr, joinEv[kSideStreams] = {};
    bool uploadPending = false;
    // all work on the engine stream (no side-stream fork for the radius classes): the request
    // server's slots overlap whole requests instead
    bool oneStream = false;
    // launch shapes (pbs_engine_set_tuning): superposition grid cap per radius class and the
    // transfer's voxel brick (index into kTransferShapes)
    int tuneSuperposeGrid = 4096;
    int tuneTransferShape = 0;
    // optional per-stage timing: kStages + 1 events per beam, resolved by pbs_engine_stage_times
    bool profiling = false;
    std::vector<hipEvent_t> evPool;
    std::vector<std::vector<hipEvent_t>> evBeams;
    // final-dose binding
    std::vector<float> fdEnergies, fdSigmas;
    float fdRaySpacing = 0.5f;
    // adjoint with respect to the spot weights: dJ/dIDD x IDD / W per (layer, step, ray), dJ/dW,
    // the transposed convolution's intermediate and the spot-lattice gradient
    DevBuf<float> adjI, adjW, adjInterm, adjSpot;
    // fixed-point dJ/dBEV of the transposed transfer (main planes, then tail planes) and the
    // float bits of max |dJ/dDose| that fix its scale
    DevBuf<unsigned long long> adjFix;
    DevBuf<unsigned> adjMax;
    // CarbonPBS profile hooks: the profile table and the per-layer row / beam parameters
    DevBuf<float> profTable;
    DevBuf<float4> profLayer;
    ConvArgs convArgs{};
    std::vector<float> hostAdjSpot;
    // staging of a no-allocation call whose pinned staging is too small: host bytes and their
    // device copy (carved from the scratch) that k_upload reads
    std::vector<char> hostStage;
    DevBuf<char> devStage;
};

// Resident state of a served beam: everything of its first run that the spot weights do not
// change (host plan, fan geometry, per-layer host values, the post-trace beam scalars and the
// per-step minimum WEPL, the subspot table, and the unit-weight IDD / rSigmaEff field with its
// passive steps). A weight-only update then re-runs the convolution, the rescale of the unit
// field, the superposition, the merge and the transfer; the trace and the fill recurrence are
// not repeated.
struct BeamState {
    BeamPlan plan;
    FanGeom geom{};
    std::vector<LayerHost> layerHost;
    int nSub = 1, maxSub = 0, numSpots = 0;
    bool subMode = false;
    DevBuf<float2> unitIdd;
    DevBuf<int> unitFp;
    DevBuf<BeamScalars> bs0;
    DevBuf<float> weplMin, subspots;
    // HIP graph of the weight-only update chain (engine-level resident beams with graphs on).
    // It is replayed only on the engine, dose buffer, launch shapes and device buffers it was
    // captured with; its uploads read the pinned copies below, which the host refills before
    // each replay once the previous replay has read them (graphDone).
    bool useGraph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void *graphEngine = nullptr;
    float *graphDose = nullptr;
    int graphDoseDims[3] = {0, 0, 0};
    int graphTune[2] = {0, 0};
    unsigned long long graphGen = 0;
    hipEvent_t graphDone = nullptr;
    float *pinWeights = nullptr;
    LayerHost *pinLayers = nullptr;
    size_t pinWeightsN = 0, pinLayersN = 0;
    void drop_graph() {
        if (graphDone) (void)hipEventSynchronize(graphDone);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graphEngine = nullptr;
        graphDose = nullptr;
    }
    BeamState() = default;
    BeamState(const BeamState &) = delete;
    BeamState &operator=(const BeamState &) = delete;
    ~BeamState() {
        drop_graph();
        if (graphDone) (void)hipEventDestroy(graphDone);
        if (pinWeights) (void)hipHostFree(pinWeights);
        if (pinLayers) (void)hipHostFree(pinLayers);
    }
    size_t bytes() const {
        return unitIdd.n * sizeof(float2) + unitFp.n * sizeof(int) + bs0.n * sizeof(BeamScalars) +
               (weplMin.n + subspots.n) * sizeof(float);
    }
};

namespace {

constexpr int kStages = 8;  // trace, conv, fill, tiles, superpose, merge, transfer, halo

hipEvent_t take_event(pbs_engine *e) {
    if (!e->evPool.empty()) {
        hipEvent_t ev = e->evPool.back();
        e->evPool.pop_back();
        return ev;
    }
    hipEvent_t ev = nullptr;
    (void)hipEventCreate(&ev);
    return ev;
}

void fill_geom(const BeamPlan &p, const pbs_engine &e, FanGeom &g) {
    std::memset(&g, 0, sizeof(g));
    const M3 *ms[3] = {&p.g2im.m, &p.dose2g.m, &p.g2dose.m};
    float *dst[3] = {g.g2im, g.d2g, g.g2dose};
    const F3 *vs[3] = {&p.g2im.v, &p.dose2g.v, &p.g2dose.v};
    float *dv[3] = {g.g2imOff, g.d2gOff, g.g2doseOff};
    for (int q = 0; q < 3; ++q) {
        const M3 &m = *ms[q];
        const float v9[9] = {m.r0.x, m.r0.y, m.r0.z, m.r1.x, m.r1.y, m.r1.z, m.r2.x, m.r2.y, m.r2.z};
        std::memcpy(dst[q], v9, sizeof(v9));
        dv[q][0] = vs[q]->x; dv[q][1] = vs[q]->y; dv[q][2] = vs[q]->z;
    }
    for (int i = 0; i < 3; ++i) { g.fc[i] = p.fanCorner[i]; g.fd[i] = p.fanDelta[i]; }
    g.dist[0] = p.sourceDist[0]; g.dist[1] = p.sourceDist[1];
    g.rx = p.rayDims[0]; g.ry = p.rayDims[1];
    g.sx = p.superpDims[0]; g.sy = p.superpDims[1];
    g.K = p.steps;
    for (int i = 0; i < 3; ++i) { g.ctDims[i] = e.ctDims[i]; g.doseDims[i] = e.doseDims[i]; }
    g.volConst = p.volConst; g.volLin = p.volLin; g.volSq = p.volSq;
    g.maxCutoffWepl = p.maxCutoffWepl;
    g.numLayers = p.numLayers;
    g.legacy = p.legacyPath ? 1 : 0;
}

// Scattering and range-energy constants of the fill: the nuclear model's eRefSq / sigmaDelta
// (idd_sigma.cu:108-121) and the table set's ion (proton defaults, idd_sigma.cu:172-181).
FillConsts fill_consts(const pbs_engine &e) {
    FillConsts fc{198.81f, 0.21f, 0, 0, 8.639415f, 0.5649718f, 938.3f, 0.0f};
    if (e.nuclearModel == 1) { fc.eRefSq = 169.00f; fc.sigmaDelta = 0.06f; fc.suppressDistalDip = 1; }  // GAUSS_FIT
    if (e.nuclearModel == 2) { fc.eRefSq = 190.44f; fc.sigmaDelta = 0.0f; }                              // SOUKUP
    if (e.nuclearModel == 3) { fc.eRefSq = 216.09f; fc.sigmaDelta = 0.08f; }                             // FLUKA
    const pbs_energy_struct &en = e.energy;
    if (en.ionRangeCoef > 0.0f) fc.rangeCoef = en.ionRangeCoef;
    if (en.ionRangeExp > 0.0f) fc.rangeExp = en.ionRangeExp;
    if (en.ionRestMass > 0.0f) fc.restMass = en.ionRestMass;
    if (en.ionScatterScale > 0.0f) fc.eRefSq *= en.ionScatterScale;
    return fc;
}

// Spot-lattice convolution arguments of a planned beam (convolution.cu:477-513).
ConvArgs make_conv_args(const BeamPlan &p, int rx, int ry, int nSub, bool subMode) {
    ConvArgs a{};
    a.rx = rx; a.ry = ry; a.spotX = p.spotDims[0]; a.spotY = p.spotDims[1];
    a.nSub = nSub; a.subspotMode = subMode ? 1 : 0;
    a.cpbResX = p.cpbRes[0]; a.cpbResY = p.cpbRes[1];
    a.deltaX = p.spotDelta[0] / p.cpbRes[0];
    a.deltaY = p.spotDelta[1] / p.cpbRes[1];
    a.offX = (p.spotOffset[0] - p.cpbCorner[0]) / p.cpbRes[0];
    a.offY = (p.spotOffset[1] - p.cpbCorner[1]) / p.cpbRes[1];
    a.spotOffX = p.spotOffset[0]; a.spotOffY = p.spotOffset[1];
    a.cpbCornerX = p.cpbCorner[0]; a.cpbCornerY = p.cpbCorner[1];
    a.lenToMm = p.lenToMm;
    return a;
}

// Host worst-case transfer launch extents: the box for steps [0, K-1] contains every
// layer's box (the fan frustum grows with its depth range).
void host_transfer_grid(const BeamPlan &p, const FanGeom &g, const int doseDims[3], int &gx, int &gy, int &gz) {
    const float xv[2] = {-32.0f, float(g.sx + 32 - 1)};
    const float yv[2] = {-32.0f, float(g.sy + 32 - 1)};
    const float zv[2] = {0.0f, float(p.steps - 1)};
    float mx[3] = {-1.f, -1.f, -1.f}, mn[3] = {1e30f, 1e30f, 1e30f};
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c) {
                F3 f = f3(xv[c] * g.fd[0] + g.fc[0], yv[b] * g.fd[1] + g.fc[1],
                          zv[a] * p.fanDelta[2] + p.fanCorner[2]);
                f.x *= (1.0f - f.z / p.sourceDist[0]);
                f.y *= (1.0f - f.z / p.sourceDist[1]);
                const F3 q = apply(p.g2dose, f);
                const float qv[3] = {q.x, q.y, q.z};
                for (int r = 0; r < 3; ++r) { mx[r] = std::max(mx[r], qv[r]); mn[r] = std::min(mn[r], qv[r]); }
            }
    const int sx = std::max((int(std::floor(mn[0])) / 32) * 32 - 32, 0);
    const int sy = std::max(int(std::floor(mn[1])) - 1, 0);
    const int ex = std::min(int(std::ceil(mx[0])) + 1, doseDims[0] - 1);
    const int ey = std::min(int(std::ceil(mx[1])) + 1, doseDims[1] - 1);
    gx = std::max(1, (ex - sx + 1 + 31) / 32 + 1);
    gy = std::max(1, (ey - sy + 1 + 7) / 8 + 1);
    const int sz = std::max(int(std::floor(mn[2])) - 1, 0);
    const int ez = std::min(int(std::ceil(mx[2])) + 1, doseDims[2] - 1);
    gz = std::max(1, ez - sz + 1);
}

void fill_transfer_args(const BeamPlan &p, const FanGeom &g, TransferArgs &t) {
    // doseIdx -> gantry (inverse of gantryToDoseIdx), then gantry -> cell index of grid g with
    // the (+32, +32) BEV halo shift; the -beamFirstInside z shift is applied on the device.
    const M3 &m = p.dose2g.m;
    const float dlt[3] = {1.0f / g.fd[0], 1.0f / g.fd[1], 1.0f / p.fanDelta[2]};
    const float col0[3] = {m.r0.x, m.r1.x, m.r2.x}, col1[3] = {m.r0.y, m.r1.y, m.r2.y}, col2[3] = {m.r0.z, m.r1.z, m.r2.z};
    const float off[3] = {p.dose2g.v.x, p.dose2g.v.y, p.dose2g.v.z};
    const float shift[3] = {32.0f, 32.0f, 0.0f};
    for (int i = 0; i < 3; ++i) {
        t.coefI[i] = col0[i] * dlt[i];
        t.coefJ[i] = col1[i] * dlt[i];
        t.inc[i] = col2[i] * dlt[i];
        t.coefOffset[i] = off[i] * dlt[i];
        const float corner = i < 2 ? g.fc[i] : p.fanCorner[2];
        t.globalOffset[i] = -corner * dlt[i] + shift[i];
    }
    t.normDist[0] = dlt[2] * p.sourceDist[0];
    t.normDist[1] = dlt[2] * p.sourceDist[1];
}

int validate_energy(const pbs_energy_struct *en) {
    if (!en) return set_error(PBS_E_INVALID, "energyData must not be null");
    if (!(en->nEnergies > 0 && en->nEnergySamples > 0)) return set_error(PBS_E_INVALID, "energyData dimensions must be positive");
    if (!en->energiesPerU || !en->peakDepths || !en->scaleFacts || !en->ciddMatrix)
        return set_error(PBS_E_INVALID, "energy axis / CIDD arrays must not be null");
    if (!en->densityVector || !en->spVector || !en->rRlVector || en->nDensitySamples <= 0 || en->nSpSamples <= 0 ||
        en->nRRlSamples <= 0)
        return set_error(PBS_E_INVALID, "density, stopping-power, and radiation-length LUTs must not be empty");
    if (en->nDensitySamples + en->nSpSamples > 16384 || en->nRRlSamples > 16384)
        return set_error(PBS_E_UNSUPPORTED, "LUT longer than the 16384 samples kept in LDS");
    return PBS_OK;
}

int validate_beam(const pbs_beam_settings &b, size_t idx) {
    const std::string pre = "beamIdx=" + std::to_string(idx) + " ";
    if (b.numLayers <= 0 || !b.energies) return set_error(PBS_E_INVALID, pre + "beam.energies must not be empty");
    if (b.numLayers > kMaxLayers) return set_error(PBS_E_UNSUPPORTED, pre + "more than 256 layers");
    if (b.maxSubspotsPerLayer <= 0 || !b.subspotData)
        return set_error(PBS_E_INVALID, pre + "beam.maxSubspotsPerLayer must be positive");
    if (!b.spotSigmas) return set_error(PBS_E_INVALID, pre + "beam.spotSigmas must not be null");
    if (b.layerSpotCounts) {
        long total = 0;
        for (int l = 0; l < b.numLayers; ++l) total += b.layerSpotCounts[l];
        if (b.spotWeights && total != b.numSpots) return set_error(PBS_E_INVALID, pre + "sum(layerSpotCounts) must match spotWeights size");
    }
    if (b.waterEquivalenceCount > 0 && (!b.waterEquivalence || b.waterEquivalenceCount < 9))
        return set_error(PBS_E_INVALID, pre + "water_equivalence needs its 9-value header");
    if (b.spotPositionsAreIndices && (!b.waterEquivalence || b.waterEquivalenceCount < 9))
        return set_error(PBS_E_INVALID, pre + "spotPositionsAreIndices requires a 9-value WEQ header");
    // the profile model's shapes (wrapper:1143-1148)
    if (b.profileData && b.profileDataCount > 0 && (!b.profileSetting || b.numProfileSetting < 3))
        return set_error(PBS_E_INVALID, pre + "profileData needs profileSetting = (first depth, depth step, depth count)");
    if (b.beamParaData && b.numBeamParaData % 3)
        return set_error(PBS_E_INVALID, pre + "beamParaData holds rows of 3 values: its length must be a multiple of 3");
    return PBS_OK;
}

}  // namespace

extern "C" {

const char *pbs_last_error(void) { return g_lastError.c_str(); }

long pbs_abi_sizeof(int which) {
    if (which == 0) return long(sizeof(pbs_beam_settings));
    if (which == 1) return long(sizeof(pbs_energy_struct));
    if (which == 2) return long(sizeof(pbs_plan_info));
    return -1;
}

int pbs_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) { (void)hipGetLastError(); return 0; }
    return n;
}

}  // extern "C"

// The engine bound by pbs_final_dose_bind for cudaFinalDose (whose reference signature carries
// no context). g_finalDoseMutex guards the pointer and is held for a whole cudaFinalDose call,
// so pbs_engine_destroy can clear the binding without racing a call that uses it; calls of
// cudaFinalDose are serialised, as the reference's blocking, global-device call was.
static std::mutex g_finalDoseMutex;
static pbs_engine *g_finalDoseEngine = nullptr;

// Passes of one beam's chain in the multi-beam batch (the record key is stage * 64 + the
// launch's index within the stage); a beam's records are in increasing key order.
enum : int {
    kStageTrace = 1,
    kStageUpload = 2,  // and 3
    kStageSetup = 4,
    kStageConv = 5,
    kStageFill = 6,
    kStageGrid = 7,
    kStageHaloSeed = 8,
    kStageHaloGrid = 9,
};
static void rec_stage(int st) {
    if (t_rec) t_rec->set_stage(st);
}

// An engine with one stream (oneStream) creates no side streams: the runtime maps streams onto
// a few hardware queues, and idle side streams of many engines would put their main streams on
// the same queue, where unrelated work serialises.
static pbs_engine *engine_create(int gpuId, const pbs_energy_struct *en, bool oneStream);

// Events that only order work (side streams against a main stream, lanes against a batch) or
// tell the host that the GPU has read pinned staging: none of them has to make device memory
// visible to the host, so they skip the system-scope fence (a cache writeback and invalidate per
// record, ~3-4 us of idle GPU each on this part). Events whose completion hands results to the
// caller keep it.
static constexpr unsigned kOrderEventFlags = hipEventDisableTiming | hipEventDisableSystemFence;

extern "C" {

pbs_engine *pbs_engine_create(int gpuId, const pbs_energy_struct *en) { return engine_create(gpuId, en, false); }

}  // extern "C"

static pbs_engine *engine_create(int gpuId, const pbs_energy_struct *en, bool oneStream) {
    if (validate_energy(en) != PBS_OK) return nullptr;
    if (hipSetDevice(gpuId) != hipSuccess) { set_error(PBS_E_DEVICE, "hipSetDevice failed"); return nullptr; }
    std::unique_ptr<pbs_engine> e(new pbs_engine());
    e->device = gpuId;
    e->oneStream = oneStream;
    (void)scratch_on_device(gpuId);  // set-up of the no-allocation entry point on this device
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&e->uploadDone, kOrderEventFlags) != hipSuccess ||
        hipEventCreateWithFlags(&e->forkEv, kOrderEventFlags) != hipSuccess) {
        set_error(PBS_E_DEVICE, "hipStreamCreate failed");
        return nullptr;
    }
    if (e->haloErr.ensure(1) != hipSuccess || hipMemset(e->haloErr.p, 0, sizeof(int)) != hipSuccess) {
        set_error(PBS_E_NOMEM, "hipMalloc failed");
        return nullptr;
    }
    for (int i = 0; i < (oneStream ? 0 : kSideStreams); ++i) {
        if (hipStreamCreateWithFlags(&e->side[i], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&e->joinEv[i], kOrderEventFlags) != hipSuccess) {
            set_error(PBS_E_DEVICE, "hipStreamCreate failed");
            return nullptr;
        }
    }
    const size_t nC = size_t(en->nEnergies) * en->nEnergySamples;
    auto up = [&](DevBuf<float> &d, const float *h, size_t n) {
        if (d.ensure(n) != hipSuccess) return false;
        return hipMemcpy(d.p, h, n * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
    };
    if (!up(e->cidd, en->ciddMatrix, nC) || !up(e->dens, en->densityVector, en->nDensitySamples) ||
        !up(e->sp, en->spVector, en->nSpSamples) || !up(e->rrl, en->rRlVector, en->nRRlSamples)) {
        set_error(PBS_E_NOMEM, "LUT upload failed");
        return nullptr;
    }
    if (en->nucWeightMatrix && en->nucSqSigmaMatrix) {
        if (!up(e->nucW, en->nucWeightMatrix, nC) || !up(e->nucS2, en->nucSqSigmaMatrix, nC)) {
            set_error(PBS_E_NOMEM, "nuclear LUT upload failed");
            return nullptr;
        }
    }
    e->hEnergies.assign(en->energiesPerU, en->energiesPerU + en->nEnergies);
    e->hPeaks.assign(en->peakDepths, en->peakDepths + en->nEnergies);
    e->hScales.assign(en->scaleFacts, en->scaleFacts + en->nEnergies);
    e->hDens.assign(en->densityVector, en->densityVector + en->nDensitySamples);
    e->hSp.assign(en->spVector, en->spVector + en->nSpSamples);
    e->env = read_plan_env();
    e->energy = *en;
    e->energy.energiesPerU = e->hEnergies.data();
    e->energy.peakDepths = e->hPeaks.data();
    e->energy.scaleFacts = e->hScales.data();
    e->energy.densityVector = e->hDens.data();
    e->energy.spVector = e->hSp.data();
    e->energy.ciddMatrix = nullptr;  // device copy only
    e->lut = LutDev{e->cidd.p, en->nEnergies, en->nEnergySamples, e->dens.p, en->nDensitySamples, en->densityScaleFact,
                    e->sp.p, en->nSpSamples, en->spScaleFact, e->rrl.p, en->nRRlSamples, en->rRlScaleFact};
    return e.release();
}

// CT adapter (wrapper:2497-2521, 2378-2441): the input type forced by RTD_CT_INPUT_TYPE or detected
// from sampled values; HU, density and SPR volumes are converted to HU+1000 on the host (density and
// SPR through the engine's density / SP table), an HU+1000 volume is used as given.
static int adapt_ct_with(const PlanEnv &env, const pbs_energy_struct &en, const float *src, size_t n, float *conv,
                         int *typeUsed) {
    const int type = env.ctInputType >= 0 ? env.ctInputType : detect_ct_type(src, n);
    if (typeUsed) *typeUsed = type;
    if (type == 1) return PBS_OK;
    const bool dens = type == 2;
    if (!ct_to_hu_plus_1000(src, n, type, dens ? en.densityVector : en.spVector,
                            dens ? en.nDensitySamples : en.nSpSamples, dens ? en.densityScaleFact : en.spScaleFact, conv))
        return set_error(PBS_E_INVALID, "CT conversion to HU+1000 failed: the density / SP table is empty or its "
                                        "scale is 0 (RTD_CT_INPUT_TYPE=HU|HU+1000|DENSITY|SPR overrides the detection)");
    return PBS_OK;
}

static int adapt_ct(const pbs_engine *e, const float *src, size_t n, std::vector<float> &conv, const float *&out) {
    conv.resize(n);
    int type = 1;
    const int rc = adapt_ct_with(e->env, e->energy, src, n, conv.data(), &type);
    if (rc != PBS_OK) return rc;
    out = type == 1 ? src : conv.data();
    if (type == 1) conv.clear();
    return PBS_OK;
}

extern "C" {

void pbs_engine_destroy(pbs_engine *e) {
    if (!e) return;
    {
        std::lock_guard<std::mutex> lk(g_finalDoseMutex);
        if (g_finalDoseEngine == e) g_finalDoseEngine = nullptr;
    }
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    if (e->uploadDone) (void)hipEventDestroy(e->uploadDone);
    for (int i = 0; i < kSideStreams; ++i) {
        if (e->side[i]) {
            (void)hipStreamSynchronize(e->side[i]);
            (void)hipStreamDestroy(e->side[i]);
        }
        if (e->joinEv[i]) (void)hipEventDestroy(e->joinEv[i]);
    }
    if (e->forkEv) (void)hipEventDestroy(e->forkEv);
    for (auto &m : e->evBeams)
        for (hipEvent_t ev : m) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : e->evPool) (void)hipEventDestroy(ev);
    if (e->pin) (void)hipHostFree(e->pin);
    delete e;
}

void *pbs_engine_stream(pbs_engine *e) { return e ? reinterpret_cast<void *>(e->stream) : nullptr; }

int pbs_engine_sync(pbs_engine *e) {
    if (!e) return set_error(PBS_E_INVALID, "null engine");
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (e->haloUsed) {
        // halo radius overflow is an error (complete_superposition.cu:558-620), raised here
        e->haloUsed = false;
        int flag = 0;
        HIP_TRY(hipMemcpy(&flag, e->haloErr.p, sizeof(int), hipMemcpyDeviceToHost));
        if (flag) {
            const int zero = 0;
            HIP_TRY(hipMemcpy(e->haloErr.p, &zero, sizeof(int), hipMemcpyHostToDevice));
            return set_error(PBS_E_GEOMETRY, "nuclear superposition kernel radius exceeded 32: the halo of a queued "
                                             "beam is wider than the largest superposition radius; its dose is invalid");
        }
    }
    return PBS_OK;
}

int pbs_engine_set_nuclear(pbs_engine *e, int model) {
    if (!e) return set_error(PBS_E_INVALID, "null engine");
    if (model < 0 || model > 3) return set_error(PBS_E_INVALID, "nuclear model must be 0 (off), 1 (GAUSS_FIT), 2 (SOUKUP) or 3 (FLUKA)");
    if (model && !(e->nucW.p && e->nucS2.p))
        return set_error(PBS_E_GEOMETRY, "nuclear_correction=true requested, but the energy struct does not carry nuclear tables");
    e->nuclearModel = model;
    return PBS_OK;
}

int pbs_engine_debug_halo(pbs_engine *e, int nucDims[2], float *spotDist, float *haloIddSigma, int *haloRadHist,
                          int *radOverflow) {
    if (!e || !e->haveLast) return set_error(PBS_E_INVALID, "no beam has been computed");
    if (!e->lastHalo) return set_error(PBS_E_INVALID, "the last beam ran without the halo branch");
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipStreamSynchronize(e->stream));
    const HaloPlan &hp = e->haloPlan;
    const int L = e->plan.numLayers, K = e->plan.steps;
    if (nucDims) { nucDims[0] = hp.nucDims[0]; nucDims[1] = hp.nucDims[1]; }
    if (spotDist) *spotDist = hp.spotDist;
    const size_t nucPlane = size_t(hp.nucDims[0]) * hp.nucDims[1];
    if (haloIddSigma)
        HIP_TRY(hipMemcpy(haloIddSigma, e->haloIddSigma.p, nucPlane * K * L * 8, hipMemcpyDeviceToHost));
    if (haloRadHist)
        HIP_TRY(hipMemcpy(haloRadHist, e->halo.hist.p, size_t(L) * (kMaxRadDev + 2) * 4, hipMemcpyDeviceToHost));
    if (radOverflow) HIP_TRY(hipMemcpy(radOverflow, e->haloErr.p, sizeof(int), hipMemcpyDeviceToHost));
    return PBS_OK;
}

int pbs_plan_debug(const int ctDims[3], const float ctRes[3], const float ctCorner[3], const int doseDims[3],
                   const float doseRes[3], const float doseCorner[3], const pbs_beam_settings *beam,
                   const pbs_energy_struct *energy, pbs_plan_info *info, float *layerEnergyIdx,
                   float *layerPeakDepth, float *denseWeights, size_t denseCap, int *spotCells,
                   float *layerProfile) {
    if (!ctDims || !ctRes || !ctCorner || !doseDims || !doseRes || !doseCorner || !beam || !energy || !info)
        return set_error(PBS_E_INVALID, "null argument");
    int rc = validate_beam(*beam, 0);
    if (rc != PBS_OK) return rc;
    BeamPlan p;
    const PlanEnv env = read_plan_env();
    if (!plan_beam(ctDims, ctRes, ctCorner, doseDims, doseRes, doseCorner, *beam, *energy, p, nullptr, &env))
        return set_error(PBS_E_GEOMETRY, p.error);
    pbs_plan_info &o = *info;
    o.legacyPath = p.legacyPath ? 1 : 0;
    o.spotIndices = p.spotIndices ? 1 : 0;
    o.weqVolume = p.weqVolume ? 1 : 0;
    o.numLayers = p.numLayers;
    o.steps = p.steps;
    for (int i = 0; i < 2; ++i) {
        o.rayDims[i] = p.rayDims[i];
        o.superpDims[i] = p.superpDims[i];
        o.spotDims[i] = p.spotDims[i];
        o.spotOffset[i] = p.spotOffset[i];
        o.spotDelta[i] = p.spotDelta[i];
        o.cpbCorner[i] = p.cpbCorner[i];
        o.cpbRes[i] = p.cpbRes[i];
        o.sourceDist[i] = p.sourceDist[i];
    }
    for (int i = 0; i < 3; ++i) {
        o.fanCorner[i] = p.fanCorner[i];
        o.fanDelta[i] = p.fanDelta[i];
    }
    o.volConst = p.volConst;
    o.volLin = p.volLin;
    o.volSq = p.volSq;
    o.lenToMm = p.lenToMm;
    o.energyDepthToMm = p.energyDepthToMm;
    o.profileHooks = p.profileHooks;
    o.profileRows = p.profileRows;
    o.profileDepthN = p.profileDepthN;
    o.profileChannels = p.profileChannels;
    if (layerProfile)
        for (size_t i = 0; i < size_t(p.numLayers) * 4; ++i)
            layerProfile[i] = i < p.profileLayer.size() ? p.profileLayer[i] : 0.0f;
    for (int l = 0; l < p.numLayers; ++l) {
        if (layerEnergyIdx) layerEnergyIdx[l] = p.layers[l].energyIdx;
        if (layerPeakDepth) layerPeakDepth[l] = p.layers[l].peakDepth;
    }
    if (denseWeights) {
        if (denseCap < p.denseWeights.size()) return set_error(PBS_E_INVALID, "denseCap is smaller than the lattice");
        std::copy(p.denseWeights.begin(), p.denseWeights.end(), denseWeights);
    }
    if (spotCells)
        for (int i = 0; i < beam->numSpots; ++i) spotCells[i] = i < int(p.spotCell.size()) ? p.spotCell[i] : -1;
    return PBS_OK;
}

int pbs_ct_to_hu_plus_1000(const float *ct, size_t n, const pbs_energy_struct *energy, float *out, int *typeUsed) {
    if (!ct || !energy || !out || n == 0) return set_error(PBS_E_INVALID, "null or empty argument");
    int type = 1;
    const int rc = adapt_ct_with(read_plan_env(), *energy, ct, n, out, &type);
    if (rc != PBS_OK) return rc;
    if (type == 1) std::memcpy(out, ct, n * sizeof(float));
    if (typeUsed) *typeUsed = type;
    return PBS_OK;
}

int pbs_engine_debug_bins(pbs_engine *e, int *effRad, int *haloEffRad) {
    if (!e || !e->haveLast) return set_error(PBS_E_INVALID, "no beam has been computed");
    if (haloEffRad && !e->lastHalo) return set_error(PBS_E_INVALID, "the last beam ran without the halo branch");
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipStreamSynchronize(e->stream));
    const size_t n = size_t(e->plan.numLayers) * (kMaxRadDev + 2);
    if (effRad) HIP_TRY(hipMemcpy(effRad, e->prim.effRad.p, n * sizeof(int), hipMemcpyDeviceToHost));
    if (haloEffRad) HIP_TRY(hipMemcpy(haloEffRad, e->halo.effRad.p, n * sizeof(int), hipMemcpyDeviceToHost));
    return PBS_OK;
}

int pbs_engine_set_ct(pbs_engine *e, const float *ct, const int dims[3], const float res[3], const float corner[3],
                      int ctIsDevice) {
    if (!e || !ct || !dims || !res || !corner) return set_error(PBS_E_INVALID, "ctData must not be null");
    if (dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0) return set_error(PBS_E_INVALID, "ctDims must be positive");
    if (!(res[0] > 0 && res[1] > 0 && res[2] > 0)) return set_error(PBS_E_INVALID, "ctResolution must be positive");
    HIP_TRY(hipSetDevice(e->device));
    const size_t n = size_t(dims[0]) * dims[1] * dims[2];
    if (ctIsDevice == 2) {  // borrowed device HU+1000 volume
        e->ctPtr = ct;
        for (int i = 0; i < 3; ++i) { e->ctDims[i] = dims[i]; e->ctRes[i] = res[i]; e->ctCorner[i] = corner[i]; }
        return PBS_OK;
    }
    std::vector<float> host;
    const float *src = ct;
    if (ctIsDevice) {
        host.resize(n);
        HIP_TRY(hipMemcpy(host.data(), ct, n * sizeof(float), hipMemcpyDeviceToHost));
        src = host.data();
    }
    std::vector<float> conv;
    int rc = adapt_ct(e, src, n, conv, src);
    if (rc != PBS_OK) return rc;
    HIP_TRY(e->ct.ensure(n));
    HIP_TRY(hipMemcpy(e->ct.p, src, n * sizeof(float), hipMemcpyHostToDevice));
    e->ctPtr = e->ct.p;
    for (int i = 0; i < 3; ++i) { e->ctDims[i] = dims[i]; e->ctRes[i] = res[i]; e->ctCorner[i] = corner[i]; }
    return PBS_OK;
}

int pbs_engine_set_dose_grid(pbs_engine *e, const int dims[3], const float res[3], const float corner[3]) {
    if (!e || !dims || !res || !corner) return set_error(PBS_E_INVALID, "null argument");
    if (dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0) return set_error(PBS_E_INVALID, "doseDims must be positive");
    if (!(res[0] > 0 && res[1] > 0 && res[2] > 0)) return set_error(PBS_E_INVALID, "doseResolution must be positive");
    for (int i = 0; i < 3; ++i) { e->doseDims[i] = dims[i]; e->doseRes[i] = res[i]; e->doseCorner[i] = corner[i]; }
    return PBS_OK;
}

static int engine_beam_dose(pbs_engine *e, const pbs_beam_settings *beam, float *doseDev, const CpbOverride *ov,
                            const float *cpbWorldGeom, const float *gradDose = nullptr, float *gradSpots = nullptr,
                            BeamState *keep = nullptr, BeamPlan *planned = nullptr, bool mayDefer = true);

int pbs_engine_beam_dose(pbs_engine *e, const pbs_beam_settings *beam, float *doseDev) {
    return engine_beam_dose(e, beam, doseDev, nullptr, nullptr);
}

int pbs_engine_beam_dose_noalloc(pbs_engine *e, const pbs_beam_settings *beam, float *doseDev) {
    if (!e) return set_error(PBS_E_INVALID, "null engine");
    if (e->profiling)
        return set_error(PBS_E_INVALID, "pbs_engine_beam_dose_noalloc: stage profiling creates events; turn it off");
    unsigned char *base = scratch_on_device(e->device);
    if (!base) return set_error(PBS_E_DEVICE, "pbs_engine_beam_dose_noalloc: the device scratch is not available");
    Arena a{base, kNoAllocScratchBytes, 0, false};
    t_arena = &a;
    const int rc = engine_beam_dose(e, beam, doseDev, nullptr, nullptr);
    t_arena = nullptr;
    if (a.overflow)
        return set_error(PBS_E_NOMEM, "pbs_engine_beam_dose_noalloc: the beam's work buffers exceed the " +
                                          std::to_string(kNoAllocScratchBytes >> 20) +
                                          " MiB no-allocation scratch; use pbs_engine_beam_dose");
    return rc;
}

int pbs_engine_beam_adjoint(pbs_engine *e, const pbs_beam_settings *beam, const float *gradDoseDev,
                            float *gradSpotWeights) {
    if (!gradDoseDev || !gradSpotWeights) return set_error(PBS_E_INVALID, "null argument");
    return engine_beam_dose(e, beam, nullptr, nullptr, nullptr, gradDoseDev, gradSpotWeights);
}

// One grid's stages after the fill: layer groups and boxes, tile radii, bins, work units;
// the radius classes on the engine and side streams; merge; transfer into the dose.
// `mark` (optional) records a stage boundary on the engine stream.
static int run_grid(pbs_engine *e, const BeamPlan &p, const FanGeom &g, GridWork &w, const float2 *iddSigma,
                    LayerScalars *ls, BeamScalars *bs, int *overflowFlag, bool codesFromFill, float *doseDev,
                    const hipStream_t streams[1 + kSideStreams], std::vector<hipEvent_t> *marks) {
    hipStream_t s = streams[0];
    auto mk = [&]() {
        if (!marks || !e->profiling) return hipSuccess;
        marks->push_back(take_event(e));
        return hipEventRecord(marks->back(), s);
    };
    HIP_TRY(mk());
    HIP_TRY(launch_finalize_and_tiles(g, iddSigma, ls, w.groupScal.p, bs, w.tailPlaneGroup.p, w.tailPlaneK.p,
                                      w.tailCap, w.tileRad.p, w.hist.p, w.effRad.p, w.itemR.p, w.tailItemR.p,
                                      w.classCount.p, w.classList.p, w.maxUnits, overflowFlag, codesFromFill ? 1 : 0,
                                      w.planeRM.p, w.slotLayers.p, s));
    HIP_TRY(mk());
    if (e->oneStream || t_rec) {
        const hipStream_t one[1 + kSideStreams] = {s, s, s};
        HIP_TRY(launch_superpose(g, iddSigma, w.tileRad.p, w.effRad.p, ls, w.groupScal.p, bs, w.classCount.p,
                                 w.classList.p, w.maxUnits, w.slotLayers.p, w.slots.p, w.tailSlots.p, e->tuneSuperposeGrid,
                                 one));
    } else {
        HIP_TRY(hipEventRecord(e->forkEv, s));
        for (int i = 0; i < kSideStreams; ++i) HIP_TRY(hipStreamWaitEvent(streams[1 + i], e->forkEv, 0));
        HIP_TRY(launch_superpose(g, iddSigma, w.tileRad.p, w.effRad.p, ls, w.groupScal.p, bs, w.classCount.p,
                                 w.classList.p, w.maxUnits, w.slotLayers.p, w.slots.p, w.tailSlots.p, e->tuneSuperposeGrid,
                                 streams));
        for (int i = 0; i < kSideStreams; ++i) {
            HIP_TRY(hipEventRecord(e->joinEv[i], streams[1 + i]));
            HIP_TRY(hipStreamWaitEvent(s, e->joinEv[i], 0));
        }
    }
    HIP_TRY(mk());
    HIP_TRY(launch_merge(g, bs, w.itemR.p, w.slots.p, w.bev.p, w.tailItemR.p, w.tailSlots.p, w.tailBev.p, w.tailCap,
                         w.planeRM.p, s));
    TransferArgs ta{};
    fill_transfer_args(p, g, ta);
    int gx = 1, gy = 1, gz = 1;
    host_transfer_grid(p, g, e->doseDims, gx, gy, gz);
    HIP_TRY(mk());
    // dose axis along which the fan depth changes fastest (row 2 of doseIdx -> gantry)
    int depthAxis = 0;
    for (int a = 1; a < 3; ++a)
        if (std::fabs(g.d2g[6 + a]) > std::fabs(g.d2g[6 + depthAxis])) depthAxis = a;
    HIP_TRY(launch_transfer(g, ta, bs, w.groupScal.p, w.bev.p, w.tailBev.p, doseDev, gx * 32, gy * 8, gz, depthAxis,
                            e->tuneTransferShape, s));
    return PBS_OK;
}

// The adjoint of one grid's stages after the fill: the same layer groups, tile radii, bins and
// work units as the forward, then the transposed transfer (into the BEV buffers, which here hold
// dJ/dBEV), the transposed superposition and IDD scaling into dJ/dW (e->adjW).
static int run_grid_adjoint(pbs_engine *e, const BeamPlan &p, const FanGeom &g, GridWork &w, const float2 *iddSigma,
                            LayerScalars *ls, BeamScalars *bs, const float *gradDose, hipStream_t s) {
    HIP_TRY(launch_finalize_and_tiles(g, iddSigma, ls, w.groupScal.p, bs, w.tailPlaneGroup.p, w.tailPlaneK.p,
                                      w.tailCap, w.tileRad.p, w.hist.p, w.effRad.p, w.itemR.p, w.tailItemR.p,
                                      w.classCount.p, w.classList.p, w.maxUnits, nullptr, 1, w.planeRM.p,
                                      w.slotLayers.p, s));
    const size_t bevPlane = size_t(g.sx + 64) * (g.sy + 64);
    const size_t mainN = size_t(g.K) * bevPlane, tailN = size_t(w.tailCap) * bevPlane;
    HIP_TRY(e->adjFix.ensure(mainN + tailN));
    HIP_TRY(e->adjMax.ensure(1));
    TransferArgs ta{};
    fill_transfer_args(p, g, ta);
    int gx = 1, gy = 1, gz = 1;
    host_transfer_grid(p, g, e->doseDims, gx, gy, gz);
    int depthAxis = 0;
    for (int a = 1; a < 3; ++a)
        if (std::fabs(g.d2g[6 + a]) > std::fabs(g.d2g[6 + depthAxis])) depthAxis = a;
    const size_t doseN = size_t(e->doseDims[0]) * e->doseDims[1] * e->doseDims[2];
    HIP_TRY(launch_adj_transfer(g, ta, bs, w.groupScal.p, gradDose, doseN, w.bev.p, w.tailBev.p, mainN, tailN,
                                e->adjFix.p, e->adjMax.p, gx * 32, gy * 8, gz, depthAxis, s));
    const size_t nray = size_t(g.rx) * g.ry;
    HIP_TRY(e->adjI.ensure(size_t(g.numLayers) * g.K * nray));
    HIP_TRY(e->adjW.ensure(size_t(g.numLayers) * nray));
    HIP_TRY(launch_adj_zero(e->adjI.p, size_t(g.numLayers) * g.K * nray, s));
    HIP_TRY(launch_adj_superpose(g, iddSigma, e->rayW.p, w.groupScal.p, bs, w.classCount.p, w.classList.p,
                                 w.maxUnits, w.slotLayers.p, w.bev.p, w.tailBev.p, e->adjI.p, s));
    HIP_TRY(launch_adj_reduce(g, ls, bs, e->adjI.p, e->adjW.p, s));
    return PBS_OK;
}

// Pinned staging of at least `bytes` (the engine's earlier uploads from it have completed: callers
// wait for uploadDone first).
static hipError_t ensure_pin(pbs_engine *e, size_t bytes) {
    if (bytes <= e->pinBytes && e->pin) return hipSuccess;
    if (e->pin) (void)hipHostFree(e->pin);
    e->pin = nullptr;
    e->pinDev = nullptr;
    e->pinBytes = 0;
    const size_t n = std::max<size_t>(bytes, size_t(1) << 20);
    hipError_t r = hipHostMalloc(reinterpret_cast<void **>(&e->pin), n);
    if (r != hipSuccess) return r;
    r = hipHostGetDevicePointer(reinterpret_cast<void **>(&e->pinDev), e->pin, 0);
    if (r != hipSuccess) {
        (void)hipHostFree(e->pin);
        e->pin = nullptr;
        e->pinDev = nullptr;
        return r;
    }
    e->pinBytes = n;
    return hipSuccess;
}

// One segment of the beam's upload: `bytes` (a multiple of 4) from the staging at `off`.
static void add_upload(UploadArgs &a, const char *stageDev, void *dst, size_t off, size_t bytes) {
    if (bytes == 0) return;
    UploadSeg &u = a.seg[a.n++];
    u.src = reinterpret_cast<const uint32_t *>(stageDev + off);
    u.dst = static_cast<uint32_t *>(dst);
    u.words = uint32_t(bytes / 4);
}

static size_t pin_align(size_t n) { return (n + 255) & ~size_t(255); }

// cpbWorldGeom (optional, 13 floats): beam direction, bmX, bmY, source, SAD of the subspot
// positions when the CPB grid is given in world coordinates (cudaFinalDose).
// `planned`: the beam's plan, made by the caller on this engine's CT, dose grid, tables and
// environment (the server plans at admission to size the request), moved in instead of planned
// again. `mayDefer`: a plain lattice beam is planned from its lattice's axes, its trace queued,
// and its spots placed while the trace runs (the placement is most of the planning time).
static int engine_beam_dose(pbs_engine *e, const pbs_beam_settings *beam, float *doseDev, const CpbOverride *ov,
                            const float *cpbWorldGeom, const float *gradDose, float *gradSpots, BeamState *keep,
                            BeamPlan *planned, bool mayDefer) {
    const bool adjoint = gradDose != nullptr;
    if (!e || !beam || (!doseDev && !adjoint)) return set_error(PBS_E_INVALID, "null argument");
    if (!e->ctPtr) return set_error(PBS_E_INVALID, "no CT bound to the engine");
    if (e->doseDims[0] <= 0) return set_error(PBS_E_INVALID, "no dose grid bound to the engine");
    int rc = validate_beam(*beam, 0);
    if (rc != PBS_OK) return rc;
    HIP_TRY(hipSetDevice(e->device));
    if (e->uploadPending) {
        HIP_TRY(hipEventSynchronize(e->uploadDone));
        e->uploadPending = false;
    }
    BeamPlan &p = e->plan;
    const bool defer = mayDefer && !planned && !ov && !cpbWorldGeom && !adjoint && !keep && e->nuclearModel == 0 &&
                       !t_rec && !t_arena;
    if (planned)
        p = std::move(*planned);
    else if (!plan_beam(e->ctDims, e->ctRes, e->ctCorner, e->doseDims, e->doseRes, e->doseCorner, *beam, e->energy, p,
                        ov, &e->env, defer))
        return set_error(PBS_E_GEOMETRY, p.error);
    FanGeom &g = e->geom;
    fill_geom(p, *e, g);
    const int L = p.numLayers, K = p.steps, rx = g.rx, ry = g.ry, sx = g.sx, sy = g.sy;
    const size_t nray = size_t(rx) * ry, nsup = size_t(sx) * sy;
    const bool subMode = p.subspotConv && !p.legacyPath;
    const int nSub = subMode ? beam->maxSubspotsPerLayer : 1;
    const bool halo = e->nuclearModel != 0;
    if (adjoint && (halo || p.legacyPath))
        return set_error(PBS_E_INVALID, halo ? "the adjoint does not cover the nuclear halo branch"
                                             : "the adjoint needs spot positions and weights (spot-lattice path)");
    if ((adjoint || keep) && p.spotIndices)
        return set_error(PBS_E_INVALID, "the adjoint and resident beams need physical spot positions, not WEQ indices");
    if (keep && (halo || p.legacyPath))
        return set_error(PBS_E_INVALID, halo ? "resident beams do not cover the nuclear halo branch"
                                             : "resident beams need spot positions and weights (spot-lattice path)");
    const bool hooks = p.profileHooks != 0;
    if (t_rec && (adjoint || keep || hooks || p.weqVolume || p.legacyPath || t_arena))
        return set_error(PBS_E_UNSUPPORTED, "the multi-beam batch covers spot-lattice beams on a CT (no adjoint, "
                                            "resident state, profile hooks, WEQ volume or legacy CPB path)");
    if (hooks && (adjoint || keep))
        return set_error(PBS_E_INVALID, "the CarbonPBS profile hooks are not covered by the adjoint or resident beams");
    if (halo) {
        if (!e->nucW.p || !e->nucS2.p)
            return set_error(PBS_E_GEOMETRY,
                             "nuclear_correction=true requested, but the energy struct does not carry nuclear tables");
        std::string herr;
        if (!plan_halo(p, *beam, e->haloPlan, herr)) return set_error(PBS_E_GEOMETRY, herr);
    }
    // buffers
    HIP_TRY(e->density.ensure(nray * K));
    HIP_TRY(e->traceScratch.ensure(trace_scratch_floats(g)));
    HIP_TRY(e->wepl.ensure(nray * K));
    HIP_TRY(e->fi.ensure(nray));
    HIP_TRY(e->fo.ensure(nray));
    HIP_TRY(e->weplMin.ensure(K));
    HIP_TRY(e->rayW.ensure(nray * L));
    HIP_TRY(e->iddSigma.ensure(nsup * K * L));
    HIP_TRY(e->firstPassive.ensure(nray * L));
    HIP_TRY(e->prim.ensure(L, K, sx, sy));
    HIP_TRY(e->layerHost.ensure(L));
    HIP_TRY(e->layerScal.ensure(L));
    HIP_TRY(e->beamScal.ensure(1));
    HIP_TRY(e->subspots.ensure(size_t(L) * beam->maxSubspotsPerLayer * 5));
    HIP_TRY(e->subEntry.ensure(size_t(L) * beam->maxSubspotsPerLayer * 2));
    FanGeom &gh = e->haloGeom;
    HaloFill hf{};
    if (halo) {
        const HaloPlan &hp = e->haloPlan;
        gh = g;  // the halo grid: cells are spots (nucIdxToFan, wrapper:5527-5535)
        gh.sx = gh.rx = hp.nucDims[0];
        gh.sy = gh.ry = hp.nucDims[1];
        gh.fc[0] = hp.offset[0]; gh.fc[1] = hp.offset[1];
        gh.fd[0] = hp.delta[0]; gh.fd[1] = hp.delta[1];
        const size_t nucPlane = size_t(hp.nucDims[0]) * hp.nucDims[1];
        HIP_TRY(e->halo.ensure(L, K, gh.sx, gh.sy));
        HIP_TRY(e->rayToNuc.ensure(nray));
        HIP_TRY(e->haloW.ensure(nucPlane * L));
        HIP_TRY(e->haloIddSigma.ensure(nucPlane * K * L));
        HIP_TRY(e->haloLayerScal.ensure(L));
        HIP_TRY(e->haloBeamScal.ensure(1));
        hf.rayToNuc = e->rayToNuc.p;
        hf.weights = e->haloW.p;
        hf.out = e->haloIddSigma.p;
        hf.nucWeightLut = e->nucW.p;
        hf.nucSqSigmaLut = e->nucS2.p;
        hf.nucPlane = int(nucPlane);
        hf.spotDist = hp.spotDist;
        hf.rSpotDistSq = 1.0f / (hp.spotDist * hp.spotDist);
    }
    std::vector<LayerHost> &lh = e->hostLayers;
    lh.resize(L);
    for (int l = 0; l < L; ++l) {
        const LayerParams &q = p.layers[l];
        lh[l] = LayerHost{q.energyIdx, q.scaleFact, q.peakDepth, q.rangeStop, q.cutoffWepl, q.airQuad, q.airLin,
                          q.convA, q.convB, q.sigmaX, q.sigmaY, 0.0f, q.airA, q.airC, 0.0f, 0.0f};
    }
    hipStream_t s = e->stream;
    std::vector<hipEvent_t> marks;
    auto mark = [&]() {
        if (!e->profiling) return hipSuccess;
        marks.push_back(take_event(e));
        return hipEventRecord(marks.back(), s);
    };
    // 1. trace and per-beam reductions
    HIP_TRY(mark());
    TracePartials parts{nullptr, nullptr, nullptr, 0, 0};
    if (p.weqVolume) {
        // the beam's water-equivalent samples replace the CT trace (wrapper:3566-3640)
        const size_t nb = size_t(p.weqDims[0]) * p.weqDims[1] * p.weqDims[2];
        HIP_TRY(e->weq.ensure(nb));
        HIP_TRY(hipMemcpyAsync(e->weq.p, beam->waterEquivalence + 9, sizeof(float) * nb, hipMemcpyHostToDevice, s));
        WeqArgs wa{e->weq.p, p.weqDims[0], p.weqDims[1], p.weqDims[2], std::min(K, p.weqDims[2]), p.weqStepMm,
                   p.cpbCorner[0], p.cpbCorner[1], p.cpbRes[0], p.cpbRes[1], p.weqX0, p.weqY0, p.weqDx, p.weqDy};
        HIP_TRY(launch_weq_fill(g, wa, e->density.p, e->wepl.p, e->fi.p, e->fo.p, s));
        HIP_TRY(launch_trace_reduce(g, e->wepl.p, e->fi.p, e->fo.p, e->weplMin.p, e->beamScal.p, s));
    } else {
        rec_stage(kStageTrace);
        HIP_TRY(launch_trace(g, e->ctPtr, e->lut, e->density.p, e->wepl.p, e->fi.p, e->fo.p, e->traceScratch.p,
                             e->weplMin.p, e->beamScal.p, &parts, s));
    }
    // the lattice's spots, placed while the trace runs
    if (p.placementPending && !place_plan_spots(*beam, p)) {
        // a weighted spot off the lattice: not a lattice beam, so its grids are not the ones just
        // traced. Once the trace is done, the beam is planned again in full (legacy CPB path)
        HIP_TRY(hipStreamSynchronize(s));
        for (hipEvent_t ev : marks) e->evPool.push_back(ev);
        return engine_beam_dose(e, beam, doseDev, ov, cpbWorldGeom, gradDose, gradSpots, keep, nullptr, false);
    }
    // the per-layer values and the subspot table go up behind the trace (copies from pageable
    // memory hold the host, so the trace is already queued while they run)
    const size_t bLayers = sizeof(LayerHost) * L, bSub = sizeof(float) * L * beam->maxSubspotsPerLayer * 5;
    const size_t bDense = p.legacyPath ? 0 : sizeof(float) * p.denseWeights.size();
    const size_t bRay = halo ? sizeof(int) * nray : 0, bHaloW = halo ? sizeof(float) * e->haloPlan.weights.size() : 0;
    const size_t oSub = pin_align(bLayers), oDense = oSub + pin_align(bSub), oRay = oDense + pin_align(bDense);
    const size_t oHaloW = oRay + pin_align(bRay);
    const size_t bProf = hooks ? sizeof(float) * size_t(p.profileRows) * p.profileDepthN * p.profileChannels : 0;
    const size_t bProfL = hooks ? sizeof(float) * p.profileLayer.size() : 0;
    const size_t oProf = oHaloW + pin_align(bHaloW), oProfL = oProf + pin_align(bProf);
    const size_t stageBytes = oProfL + pin_align(bProfL);
    // the staging k_upload reads: the mapped pinned buffer, or (a no-allocation call whose pinned
    // staging is too small) a device copy of host bytes carved from the scratch
    const bool devStage = t_arena && (!e->pin || e->pinBytes < stageBytes);
    char *stageHost = nullptr;
    const char *stageSrc = nullptr;
    if (devStage) {
        e->hostStage.resize(stageBytes);
        HIP_TRY(e->devStage.ensure(stageBytes));
        stageHost = e->hostStage.data();
        stageSrc = e->devStage.p;
    } else {
        HIP_TRY(ensure_pin(e, stageBytes));
        stageHost = e->pin;
        stageSrc = e->pinDev;
    }
    std::memcpy(stageHost, lh.data(), bLayers);
    std::memcpy(stageHost + oSub, beam->subspotData, bSub);
    if (bDense) std::memcpy(stageHost + oDense, p.denseWeights.data(), bDense);
    if (halo) {
        std::memcpy(stageHost + oRay, e->haloPlan.rayToNuc.data(), bRay);
        std::memcpy(stageHost + oHaloW, e->haloPlan.weights.data(), bHaloW);
    }
    if (hooks) {
        std::memcpy(stageHost + oProf, beam->profileData, bProf);
        std::memcpy(stageHost + oProfL, p.profileLayer.data(), bProfL);
    }
    if (devStage) HIP_TRY(hipMemcpyAsync(e->devStage.p, stageHost, stageBytes, hipMemcpyHostToDevice, s));
    UploadArgs up{};
    add_upload(up, stageSrc, e->layerHost.p, 0, bLayers);
    add_upload(up, stageSrc, e->subspots.p, oSub, bSub);
    if (bDense) {
        HIP_TRY(e->spot.ensure(p.denseWeights.size()));
        add_upload(up, stageSrc, e->spot.p, oDense, bDense);
    }
    if (halo) {
        add_upload(up, stageSrc, e->rayToNuc.p, oRay, bRay);
        add_upload(up, stageSrc, e->haloW.p, oHaloW, bHaloW);
        rec_stage(kStageUpload);
        HIP_TRY(launch_clear_pairs(e->haloIddSigma.p, size_t(hf.nucPlane) * K * L, s));
    }
    if (hooks) {
        HIP_TRY(e->profTable.ensure(bProf / sizeof(float)));
        HIP_TRY(e->profLayer.ensure(size_t(L)));
        add_upload(up, stageSrc, e->profTable.p, oProf, bProf);
        add_upload(up, stageSrc, e->profLayer.p, oProfL, bProfL);
    }
    // the trace's partials (one-pass trace) are reduced by the same launch
    up.tp = parts;
    up.weplMin = e->weplMin.p;
    up.bs = e->beamScal.p;
    up.K = K;
    rec_stage(kStageUpload + 1);
    HIP_TRY(launch_upload(up, s));
    rec_stage(kStageSetup);
    const float nucFactor = e->nuclearModel == 1 ? 0.97f : 1.0f;  // GAUSS_FIT entry sigma, wrapper:3788-3795
    HIP_TRY(launch_beam_setup(g, e->weplMin.p, e->layerHost.p, e->beamScal.p, e->layerScal.p,
                              subMode ? e->subspots.p : nullptr, subMode ? nSub : 0, e->subEntry.p, nucFactor,
                              e->prim.classCount.p, e->prim.hist.p, s));
    if (keep) {
        HIP_TRY(keep->bs0.ensure(1));
        HIP_TRY(keep->weplMin.ensure(K));
        HIP_TRY(keep->subspots.ensure(size_t(L) * beam->maxSubspotsPerLayer * 5));
        HIP_TRY(keep->unitIdd.ensure(nsup * K * L));
        HIP_TRY(keep->unitFp.ensure(nray * L));
        HIP_TRY(hipMemcpyAsync(keep->bs0.p, e->beamScal.p, sizeof(BeamScalars), hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(keep->weplMin.p, e->weplMin.p, sizeof(float) * K, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(keep->subspots.p, e->subspots.p, sizeof(float) * L * beam->maxSubspotsPerLayer * 5,
                               hipMemcpyDeviceToDevice, s));
        keep->plan = p;
        keep->geom = g;
        keep->layerHost = lh;
        keep->nSub = nSub;
        keep->maxSub = beam->maxSubspotsPerLayer;
        keep->subMode = subMode;
        keep->numSpots = beam->numSpots;
    }
    // 2. ray weights
    HIP_TRY(mark());
    if (!p.legacyPath) {
        // the dense spot lattice went up with the beam's other arrays (k_upload above)
        HIP_TRY(e->interm.ensure(size_t(L) * nSub * p.spotDims[1] * rx));
        const ConvArgs a = make_conv_args(p, rx, ry, nSub, subMode);
        rec_stage(kStageConv);
        HIP_TRY(launch_conv(a, e->spot.p, e->interm.p, e->rayW.p, e->layerScal.p, e->beamScal.p, e->subspots.p,
                            e->subEntry.p, L, s));
        e->convArgs = a;
    } else {
        // subspot -> CPB in the gantry frame (wrapper:3480-3511); the ray grid is the CPB grid
        CpbArgs a{};
        a.nx = rx; a.ny = ry; a.nSub = beam->maxSubspotsPerLayer;
        a.cx = p.cpbCorner[0]; a.cy = p.cpbCorner[1]; a.rx_ = p.cpbRes[0]; a.ry_ = p.cpbRes[1];
        const float bd[3] = {0.f, 0.f, -1.f}, bx[3] = {1.f, 0.f, 0.f}, by[3] = {0.f, -1.f, 0.f};
        for (int i = 0; i < 3; ++i) { a.bd[i] = bd[i]; a.bx[i] = bx[i]; a.by[i] = by[i]; }
        a.src[0] = 0.f; a.src[1] = 0.f; a.src[2] = p.sadInput;
        a.sad = p.sadInput;
        if (cpbWorldGeom) {
            for (int i = 0; i < 3; ++i) {
                a.bd[i] = cpbWorldGeom[i];
                a.bx[i] = cpbWorldGeom[3 + i];
                a.by[i] = cpbWorldGeom[6 + i];
                a.src[i] = cpbWorldGeom[9 + i];
            }
            a.sad = cpbWorldGeom[12];
        }
        HIP_TRY(e->cpb.ensure(nray * L));
        HIP_TRY(launch_cpb(a, e->subspots.p, e->cpb.p, e->rayW.p, rx, ry, p.cpbCorner[0], p.cpbCorner[1], p.cpbRes[0],
                           p.cpbRes[1], L, s));
    }
    if (!t_rec) {  // the batch executor records it after the batch's launches
        HIP_TRY(hipEventRecord(e->uploadDone, s));
        e->uploadPending = true;
    }
    // 3. IDD / sigma (+ the halo's share when the halo branch is on; idd_sigma.cu:108-121)
    HIP_TRY(mark());
    FillConsts fc = fill_consts(*e);
    if (keep) {
        // the weight-independent field (every ray at weight 1) into the resident state, then the
        // layer scalars reset and the rescale to this request's weights
        fc.unitWeight = 1;
        HIP_TRY(launch_fill(g, e->lut, e->layerHost.p, e->layerScal.p, e->beamScal.p, e->density.p, e->wepl.p,
                            e->rayW.p, e->fi.p, e->fo.p, keep->unitIdd.p, keep->unitFp.p, fc, hf, e->prim.tileRad.p,
                            s));
        HIP_TRY(launch_beam_setup(g, e->weplMin.p, e->layerHost.p, e->beamScal.p, e->layerScal.p,
                                  subMode ? e->subspots.p : nullptr, subMode ? nSub : 0, e->subEntry.p, 1.0f,
                                  e->prim.classCount.p, e->prim.hist.p, s));
        HIP_TRY(launch_rescale(g, e->layerScal.p, e->beamScal.p, e->rayW.p, keep->unitIdd.p,
                               keep->unitFp.p, e->iddSigma.p, e->firstPassive.p, e->prim.tileRad.p, s));
    } else {
        rec_stage(kStageFill);
        HIP_TRY(launch_fill(g, e->lut, e->layerHost.p, e->layerScal.p, e->beamScal.p, e->density.p, e->wepl.p,
                            e->rayW.p, e->fi.p, e->fo.p, e->iddSigma.p, e->firstPassive.p, fc, hf, e->prim.tileRad.p,
                            s));
    }
    if (hooks) {
        // the profile hooks rewrite the field after the fill (wrapper:4604-4832); the tile codes
        // are then taken from the rewritten field instead of the fill's
        ProfileArgs pa{};
        pa.table = e->profTable.p;
        pa.layer = e->profLayer.p;
        pa.rows = p.profileRows;
        pa.depthN = p.profileDepthN;
        pa.channels = p.profileChannels;
        pa.hooks = p.profileHooks;
        pa.depth0 = p.profileDepth0;
        pa.depthStep = p.profileDepthStep;
        pa.beamParaPos = p.beamParaPos;
        HIP_TRY(launch_profile_hooks(g, pa, e->wepl.p, e->iddSigma.p, e->layerScal.p, e->beamScal.p, s));
    }
    if (adjoint) {
        // 4-6 transposed, then the transposed convolution and the spots' lattice cells
        int rc3 = run_grid_adjoint(e, p, g, e->prim, e->iddSigma.p, e->layerScal.p, e->beamScal.p, gradDose, s);
        if (rc3 != PBS_OK) return rc3;
        const size_t splane = size_t(p.spotDims[0]) * p.spotDims[1];
        HIP_TRY(e->adjInterm.ensure(size_t(L) * nSub * p.spotDims[1] * rx));
        HIP_TRY(e->adjSpot.ensure(splane * L));
        HIP_TRY(launch_adj_conv(e->convArgs, e->adjW.p, e->adjInterm.p, e->adjSpot.p, e->layerScal.p, e->beamScal.p,
                                e->subspots.p, e->subEntry.p, L, s));
        e->hostAdjSpot.resize(splane * L);
        HIP_TRY(hipMemcpyAsync(e->hostAdjSpot.data(), e->adjSpot.p, sizeof(float) * splane * L,
                               hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (int i = 0; i < beam->numSpots; ++i) {
            const int c = i < int(p.spotCell.size()) ? p.spotCell[i] : -1;
            gradSpots[i] = c >= 0 ? e->hostAdjSpot[size_t(c)] : 0.0f;
        }
        e->haveLast = true;
        return PBS_OK;
    }
    // 4-6. primary grid: passive steps, tiles, bins, items; superposition; merge; transfer
    hipStream_t streams[1 + kSideStreams] = {s};
    for (int i = 0; i < kSideStreams; ++i) streams[1 + i] = e->oneStream ? s : e->side[i];
    rec_stage(kStageGrid);
    int rc2 = run_grid(e, p, g, e->prim, e->iddSigma.p, e->layerScal.p, e->beamScal.p, nullptr, !hooks, doseDev,
                       streams, &marks);
    if (rc2 != PBS_OK) return rc2;
    // 7. halo grid through the same stages
    HIP_TRY(mark());
    if (halo) {
        rec_stage(kStageHaloSeed);
        HIP_TRY(launch_seed_grid(L, e->layerScal.p, e->beamScal.p, e->haloLayerScal.p, e->haloBeamScal.p,
                                 e->halo.hist.p, e->halo.classCount.p, s));
        rec_stage(kStageHaloGrid);
        rc2 = run_grid(e, p, gh, e->halo, e->haloIddSigma.p, e->haloLayerScal.p, e->haloBeamScal.p, e->haloErr.p,
                       false, doseDev, streams, nullptr);
        if (rc2 != PBS_OK) return rc2;
    }
    e->lastHalo = halo;
    e->haloUsed |= halo;
    HIP_TRY(mark());
    if (e->profiling) e->evBeams.push_back(std::move(marks));
    e->haveLast = true;
    return PBS_OK;
}

// Weight-only update of a resident beam on engine e: the lattice re-binned with the new
// weights, the convolution, the rescale of the resident unit-weight field, then the grid
// stages into doseDev (accumulated). The CT, dose grid and nuclear model of e must be those of
// the state's first run.
// Host part of a weight-only update: the resident plan with the lattice re-binned at the new
// weights (e->plan.denseWeights), the fan geometry, and every buffer the chain needs sized.
static int update_prepare(pbs_engine *e, const BeamState &st, const float *weights, int numSpots) {
    if (!e || !weights) return set_error(PBS_E_INVALID, "null argument");
    if (numSpots != st.numSpots) return set_error(PBS_E_INVALID, "weight count differs from the resident beam's");
    if (e->nuclearModel != 0) return set_error(PBS_E_INVALID, "resident beams do not cover the nuclear halo branch");
    HIP_TRY(hipSetDevice(e->device));
    if (e->uploadPending) {
        HIP_TRY(hipEventSynchronize(e->uploadDone));
        e->uploadPending = false;
    }
    BeamPlan &p = e->plan;
    p = st.plan;
    std::fill(p.denseWeights.begin(), p.denseWeights.end(), 0.0f);
    for (int i = 0; i < numSpots; ++i) {
        const float w = weights[i];
        if (!std::isfinite(w) || w < 0.0f) return set_error(PBS_E_INVALID, "spot weights must be finite and >= 0");
        if (w == 0.0f) continue;
        const int c = p.spotCell[size_t(i)];
        if (c < 0) return set_error(PBS_E_GEOMETRY, "a weighted spot lies outside the resident beam's lattice");
        p.denseWeights[size_t(c)] += w;
    }
    FanGeom &g = e->geom;
    g = st.geom;
    for (int i = 0; i < 3; ++i) g.doseDims[i] = e->doseDims[i];
    const int L = p.numLayers, K = p.steps, rx = g.rx, ry = g.ry;
    const size_t nray = size_t(rx) * ry, nsup = size_t(g.sx) * g.sy;
    const size_t splane = size_t(p.spotDims[0]) * p.spotDims[1];
    HIP_TRY(e->weplMin.ensure(K));
    HIP_TRY(e->rayW.ensure(nray * L));
    HIP_TRY(e->iddSigma.ensure(nsup * K * L));
    HIP_TRY(e->firstPassive.ensure(nray * L));
    HIP_TRY(e->prim.ensure(L, K, g.sx, g.sy));
    HIP_TRY(e->layerHost.ensure(L));
    HIP_TRY(e->layerScal.ensure(L));
    HIP_TRY(e->beamScal.ensure(1));
    HIP_TRY(e->subEntry.ensure(size_t(L) * std::max(st.maxSub, 1) * 2));
    HIP_TRY(e->spot.ensure(splane * L));
    HIP_TRY(e->interm.ensure(size_t(L) * st.nSub * p.spotDims[1] * rx));
    return PBS_OK;
}

// Device part of a weight-only update (after update_prepare), all on the engine stream: the
// per-layer values and the re-binned lattice uploaded from wPlane / layers, the post-trace
// scalars restored, cut-offs, convolution, the rescale of the unit-weight field, then either the
// grid stages into doseDev or (gradDose given) the adjoint into gradSpots. markUpload: record
// the engine's upload event (eager calls; a captured chain reads pinned copies instead).
static int update_enqueue(pbs_engine *e, const BeamState &st, const float *wPlane, const LayerHost *layers,
                          float *doseDev, bool markUpload, const float *gradDose, float *gradSpots) {
    const BeamPlan &p = e->plan;
    const FanGeom &g = e->geom;
    const int L = p.numLayers, K = p.steps, rx = g.rx, ry = g.ry;
    const size_t splane = size_t(p.spotDims[0]) * p.spotDims[1];
    hipStream_t s = e->stream;
    HIP_TRY(hipMemcpyAsync(e->layerHost.p, layers, sizeof(LayerHost) * L, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(e->spot.p, wPlane, sizeof(float) * splane * L, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(e->beamScal.p, st.bs0.p, sizeof(BeamScalars), hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(e->weplMin.p, st.weplMin.p, sizeof(float) * K, hipMemcpyDeviceToDevice, s));
    HIP_TRY(launch_beam_setup(g, e->weplMin.p, e->layerHost.p, e->beamScal.p, e->layerScal.p,
                              st.subMode ? st.subspots.p : nullptr, st.subMode ? st.nSub : 0, e->subEntry.p, 1.0f,
                              e->prim.classCount.p, e->prim.hist.p, s));
    const ConvArgs a = make_conv_args(p, rx, ry, st.nSub, st.subMode);
    HIP_TRY(launch_conv(a, e->spot.p, e->interm.p, e->rayW.p, e->layerScal.p, e->beamScal.p, st.subspots.p,
                        e->subEntry.p, L, s));
    e->convArgs = a;
    if (markUpload) {
        HIP_TRY(hipEventRecord(e->uploadDone, s));
        e->uploadPending = true;
    }
    HIP_TRY(launch_rescale(g, e->layerScal.p, e->beamScal.p, e->rayW.p, st.unitIdd.p, st.unitFp.p,
                           e->iddSigma.p, e->firstPassive.p, e->prim.tileRad.p, s));
    if (gradDose) {
        int rc = run_grid_adjoint(e, p, g, e->prim, e->iddSigma.p, e->layerScal.p, e->beamScal.p, gradDose, s);
        if (rc != PBS_OK) return rc;
        HIP_TRY(e->adjInterm.ensure(size_t(L) * st.nSub * p.spotDims[1] * rx));
        HIP_TRY(e->adjSpot.ensure(splane * L));
        HIP_TRY(launch_adj_conv(e->convArgs, e->adjW.p, e->adjInterm.p, e->adjSpot.p, e->layerScal.p, e->beamScal.p,
                                st.subspots.p, e->subEntry.p, L, s));
        e->hostAdjSpot.resize(splane * L);
        HIP_TRY(hipMemcpyAsync(e->hostAdjSpot.data(), e->adjSpot.p, sizeof(float) * splane * L,
                               hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (int i = 0; i < st.numSpots; ++i) {
            const int c = i < int(p.spotCell.size()) ? p.spotCell[size_t(i)] : -1;
            gradSpots[i] = c >= 0 ? e->hostAdjSpot[size_t(c)] : 0.0f;
        }
    } else {
        hipStream_t streams[1 + kSideStreams] = {s};
        for (int i = 0; i < kSideStreams; ++i) streams[1 + i] = e->oneStream ? s : e->side[i];
        int rc = run_grid(e, p, g, e->prim, e->iddSigma.p, e->layerScal.p, e->beamScal.p, nullptr, true, doseDev,
                          streams, nullptr);
        if (rc != PBS_OK) return rc;
    }
    e->lastHalo = false;
    e->haveLast = true;
    return PBS_OK;
}

// Weight-only update of a resident beam on engine e: the lattice re-binned with the new
// weights, the convolution, the rescale of the resident unit-weight field, then the grid
// stages into doseDev (accumulated), or the adjoint into gradSpots when gradDose is given. The
// CT, dose grid and nuclear model of e must be those of the state's first run.
static int engine_weights_update(pbs_engine *e, const BeamState &st, const float *weights, int numSpots,
                                 float *doseDev, const float *gradDose = nullptr, float *gradSpots = nullptr) {
    if (!doseDev && !gradDose) return set_error(PBS_E_INVALID, "null argument");
    int rc = update_prepare(e, st, weights, numSpots);
    if (rc != PBS_OK) return rc;
    e->hostLayers = st.layerHost;
    return update_enqueue(e, st, e->plan.denseWeights.data(), e->hostLayers.data(), doseDev, true, gradDose,
                          gradSpots);
}

// The same update through a HIP graph of the whole chain (state.useGraph): replayed while the
// engine, dose buffer, dose grid, launch shapes and every device buffer are the ones captured;
// otherwise the update runs eagerly and the chain is captured again for the next call.
static int engine_weights_update_graph(pbs_engine *e, BeamState &st, const float *weights, int numSpots,
                                       float *doseDev) {
    if (!st.useGraph || e->profiling) return engine_weights_update(e, st, weights, numSpots, doseDev);
    if (!doseDev) return set_error(PBS_E_INVALID, "null argument");
    const bool valid = st.exec && st.graphEngine == e && st.graphDose == doseDev &&
                       st.graphDoseDims[0] == e->doseDims[0] && st.graphDoseDims[1] == e->doseDims[1] &&
                       st.graphDoseDims[2] == e->doseDims[2] && st.graphTune[0] == e->tuneSuperposeGrid &&
                       st.graphTune[1] == e->tuneTransferShape && st.graphGen == g_bufGen.load();
    if (valid) {
        int rc = update_prepare(e, st, weights, numSpots);
        if (rc != PBS_OK) return rc;
        // the previous replay's uploads have read the pinned weights before they are rewritten
        HIP_TRY(hipEventSynchronize(st.graphDone));
        std::memcpy(st.pinWeights, e->plan.denseWeights.data(), sizeof(float) * e->plan.denseWeights.size());
        HIP_TRY(hipGraphLaunch(st.exec, e->stream));
        HIP_TRY(hipEventRecord(st.graphDone, e->stream));
        e->lastHalo = false;
        e->haveLast = true;
        return PBS_OK;
    }
    // eager update now (sizes every buffer), then capture the chain for the next calls
    int rc = engine_weights_update(e, st, weights, numSpots, doseDev);
    if (rc != PBS_OK) return rc;
    st.drop_graph();
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->uploadPending = false;
    const size_t nw = e->plan.denseWeights.size(), nl = st.layerHost.size();
    if (st.pinWeightsN < nw) {
        if (st.pinWeights) HIP_TRY(hipHostFree(st.pinWeights));
        st.pinWeights = nullptr;
        st.pinWeightsN = 0;
        HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&st.pinWeights), sizeof(float) * std::max<size_t>(nw, 1)));
        st.pinWeightsN = nw;
    }
    if (st.pinLayersN < nl) {
        if (st.pinLayers) HIP_TRY(hipHostFree(st.pinLayers));
        st.pinLayers = nullptr;
        st.pinLayersN = 0;
        HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&st.pinLayers), sizeof(LayerHost) * std::max<size_t>(nl, 1)));
        st.pinLayersN = nl;
    }
    if (!st.graphDone) HIP_TRY(hipEventCreateWithFlags(&st.graphDone, hipEventDisableTiming));
    std::memcpy(st.pinWeights, e->plan.denseWeights.data(), sizeof(float) * nw);
    std::memcpy(st.pinLayers, st.layerHost.data(), sizeof(LayerHost) * nl);
    const unsigned long long gen = g_bufGen.load();
    const bool one = e->oneStream;
    e->oneStream = true;  // the radius classes on the engine stream only: one linear chain
    HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeRelaxed));
    rc = update_enqueue(e, st, st.pinWeights, st.pinLayers, doseDev, false, nullptr, nullptr);
    hipGraph_t graph = nullptr;
    const hipError_t ce = hipStreamEndCapture(e->stream, &graph);
    e->oneStream = one;
    if (rc != PBS_OK || ce != hipSuccess || g_bufGen.load() != gen) {
        if (graph) (void)hipGraphDestroy(graph);
        if (rc != PBS_OK) return rc;
        if (ce != hipSuccess) return set_error(PBS_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
        return PBS_OK;  // a buffer moved during capture: stay eager, capture again next call
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return set_error(PBS_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    }
    st.graph = graph;
    st.exec = exec;
    st.graphEngine = e;
    st.graphDose = doseDev;
    for (int i = 0; i < 3; ++i) st.graphDoseDims[i] = e->doseDims[i];
    st.graphTune[0] = e->tuneSuperposeGrid;
    st.graphTune[1] = e->tuneTransferShape;
    st.graphGen = gen;
    HIP_TRY(hipEventRecord(st.graphDone, e->stream));
    return PBS_OK;
}

struct pbs_resident_beam {
    BeamState st;
    pbs_engine *engine = nullptr;
};

int pbs_engine_keep_beam(pbs_engine *e, const pbs_beam_settings *beam, float *doseDev, pbs_resident_beam **out) {
    if (!e || !beam || !doseDev || !out) return set_error(PBS_E_INVALID, "null argument");
    *out = nullptr;
    std::unique_ptr<pbs_resident_beam> rb(new pbs_resident_beam());
    rb->engine = e;
    int rc = engine_beam_dose(e, beam, doseDev, nullptr, nullptr, nullptr, nullptr, &rb->st);
    if (rc != PBS_OK) return rc;
    *out = rb.release();
    return PBS_OK;
}

int pbs_engine_update_beam(pbs_engine *e, pbs_resident_beam *rb, const float *spotWeights, int numSpots,
                           float *doseDev) {
    if (!e || !rb || !spotWeights || !doseDev) return set_error(PBS_E_INVALID, "null argument");
    if (rb->engine != e) return set_error(PBS_E_INVALID, "the resident beam belongs to another engine");
    return engine_weights_update_graph(e, rb->st, spotWeights, numSpots, doseDev);
}

int pbs_engine_update_adjoint(pbs_engine *e, pbs_resident_beam *rb, const float *spotWeights, int numSpots,
                              const float *gradDoseDev, float *gradSpotWeights) {
    if (!e || !rb || !spotWeights || !gradDoseDev || !gradSpotWeights) return set_error(PBS_E_INVALID, "null argument");
    if (rb->engine != e) return set_error(PBS_E_INVALID, "the resident beam belongs to another engine");
    return engine_weights_update(e, rb->st, spotWeights, numSpots, nullptr, gradDoseDev, gradSpotWeights);
}

int pbs_resident_beam_use_graph(pbs_resident_beam *rb, int enable) {
    if (!rb) return set_error(PBS_E_INVALID, "null argument");
    (void)hipSetDevice(rb->engine->device);
    rb->st.useGraph = enable != 0;
    if (!rb->st.useGraph) rb->st.drop_graph();
    return PBS_OK;
}

void pbs_resident_beam_destroy(pbs_resident_beam *rb) {
    if (!rb) return;
    (void)hipSetDevice(rb->engine->device);
    (void)hipStreamSynchronize(rb->engine->stream);
    delete rb;
}

int pbs_engine_set_profiling(pbs_engine *e, int enable) {
    if (!e) return set_error(PBS_E_INVALID, "null engine");
    e->profiling = enable != 0;
    return PBS_OK;
}

int pbs_engine_set_tuning(pbs_engine *e, int key, int value) {
    if (!e) return set_error(PBS_E_INVALID, "null engine");
    switch (key) {
    case PBS_TUNE_SUPERPOSE_GRID:
        if (value < 64 || value > 65536) return set_error(PBS_E_INVALID, "superposition grid must be 64..65536");
        e->tuneSuperposeGrid = value;
        return PBS_OK;
    case PBS_TUNE_TRANSFER_SHAPE:
        if (value < 0 || value >= kNumTransferShapes) return set_error(PBS_E_INVALID, "unknown transfer shape");
        e->tuneTransferShape = value;
        return PBS_OK;
    case PBS_TUNE_ONE_STREAM:
        if (value == 0 && !e->side[0]) return set_error(PBS_E_INVALID, "this engine was created without side streams");
        e->oneStream = value != 0;
        return PBS_OK;
    default:
        return set_error(PBS_E_INVALID, "unknown tuning key");
    }
}

int pbs_engine_get_tuning(pbs_engine *e, int key, int *value) {
    if (!e || !value) return set_error(PBS_E_INVALID, "null argument");
    switch (key) {
    case PBS_TUNE_SUPERPOSE_GRID: *value = e->tuneSuperposeGrid; return PBS_OK;
    case PBS_TUNE_TRANSFER_SHAPE: *value = e->tuneTransferShape; return PBS_OK;
    case PBS_TUNE_ONE_STREAM: *value = e->oneStream ? 1 : 0; return PBS_OK;
    default: return set_error(PBS_E_INVALID, "unknown tuning key");
    }
}

int pbs_engine_stage_times(pbs_engine *e, float *msOut, int *beams) {
    if (!e) return set_error(PBS_E_INVALID, "null engine");
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipStreamSynchronize(e->stream));
    std::vector<double> acc(kStages, 0.0);
    for (auto &m : e->evBeams) {
        for (int i = 0; i < kStages && i + 1 < int(m.size()); ++i) {
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, m[i], m[i + 1]));
            acc[i] += ms;
        }
        for (hipEvent_t ev : m) e->evPool.push_back(ev);
    }
    if (beams) *beams = int(e->evBeams.size());
    e->evBeams.clear();
    if (msOut)
        for (int i = 0; i < kStages; ++i) msOut[i] = float(acc[i]);
    return PBS_OK;
}

int pbs_engine_debug_fetch(pbs_engine *e, pbs_debug_ints *ints, float *density, float *wepl, int *firstInside,
                           int *firstOutside, float *weplMin, float *rayWeights, float *iddSigma, int *firstPassive,
                           int *tileRad, int *radHist, float *bev) {
    if (!e || !e->haveLast) return set_error(PBS_E_INVALID, "no beam has been computed");
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipStreamSynchronize(e->stream));
    const BeamPlan &p = e->plan;
    const FanGeom &g = e->geom;
    const int L = p.numLayers, K = p.steps;
    const size_t nray = size_t(g.rx) * g.ry, nsup = size_t(g.sx) * g.sy;
    const int nt = (g.sx / 32) * (g.sy / 8);
    auto cp = [&](void *dst, const void *src, size_t bytes) {
        return dst ? hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) : hipSuccess;
    };
    HIP_TRY(cp(density, e->density.p, nray * K * 4));
    HIP_TRY(cp(wepl, e->wepl.p, nray * K * 4));
    HIP_TRY(cp(firstInside, e->fi.p, nray * 4));
    HIP_TRY(cp(firstOutside, e->fo.p, nray * 4));
    HIP_TRY(cp(weplMin, e->weplMin.p, size_t(K) * 4));
    HIP_TRY(cp(rayWeights, e->rayW.p, nray * L * 4));
    HIP_TRY(cp(iddSigma, e->iddSigma.p, nsup * K * L * 8));
    HIP_TRY(cp(firstPassive, e->firstPassive.p, nray * L * 4));
    HIP_TRY(cp(tileRad, e->prim.tileRad.p, size_t(L) * K * nt * 4));
    HIP_TRY(cp(radHist, e->prim.hist.p, size_t(L) * (kMaxRadDev + 2) * 4));
    HIP_TRY(cp(bev, e->prim.bev.p, size_t(K) * (g.sx + 64) * (g.sy + 64) * 4));
    if (ints) {
        BeamScalars bs{};
        std::vector<LayerScalars> ls(L);
        HIP_TRY(hipMemcpy(&bs, e->beamScal.p, sizeof(bs), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(ls.data(), e->layerScal.p, sizeof(LayerScalars) * L, hipMemcpyDeviceToHost));
        std::memset(ints, 0, sizeof(*ints));
        ints->rayDims[0] = g.rx; ints->rayDims[1] = g.ry;
        ints->superpDims[0] = g.sx; ints->superpDims[1] = g.sy;
        ints->steps = K; ints->numLayers = L; ints->legacyPath = g.legacy;
        ints->beamFirstInside = bs.beamFirstInside;
        ints->beamFirstOutside = bs.beamFirstOutside;
        ints->firstPastCutoffAll = bs.firstPastCutoffAll;
        ints->guaranteedPassive = bs.guaranteedPassive;
        for (int l = 0; l < L; ++l) {
            ints->afterLast[l] = ls[l].afterLast;
            ints->layerPassive[l] = ls[l].passive;
            ints->energyIdx[l] = p.layers[l].energyIdx;
            ints->scaleFact[l] = p.layers[l].scaleFact;
            ints->peakDepth[l] = p.layers[l].peakDepth;
        }
        for (int i = 0; i < 3; ++i) {
            ints->transferStart[i] = bs.boxStart[i];
            ints->transferMax[i] = bs.boxMax[i];
            ints->cpbCorner[i] = p.cpbCorner[i];
            ints->fanCorner[i] = p.fanCorner[i];
            ints->fanDelta[i] = p.fanDelta[i];
        }
        ints->sourceDist[0] = p.sourceDist[0];
        ints->sourceDist[1] = p.sourceDist[1];
        ints->numGroups = bs.numGroups;
        ints->tailPlanes = bs.tailPlanes;
        ints->tailOverflow = bs.overflow;
        ints->mainEndMax = bs.mainEndMax;
        std::vector<GroupScalars> gs(std::max(bs.numGroups, 1));
        if (bs.numGroups > 0)
            HIP_TRY(hipMemcpy(gs.data(), e->prim.groupScal.p, sizeof(GroupScalars) * bs.numGroups,
                              hipMemcpyDeviceToHost));
        for (int l = 0; l < L; ++l) {
            ints->mainEnd[l] = ls[l].mainEnd;
            ints->layerGroup[l] = ls[l].group;
        }
        for (int j = 0; j < bs.numGroups && j < 256; ++j) {
            ints->groupPassive[j] = gs[j].passive;
            ints->groupTailStart[j] = gs[j].tailStart;
            ints->groupTailCount[j] = gs[j].tailCount;
        }
    }
    return PBS_OK;
}

int pbs_subsecond_wrapper(const float *ctData, const int ctDims[3], const float ctResolution[3],
                          const float ctCorner[3], float *doseData, const int doseDims[3],
                          const float doseResolution[3], const float doseCorner[3], const pbs_beam_settings *beams,
                          size_t numBeams, const pbs_energy_struct *energy, int gpuId, int nuclearCorrection,
                          int verbose) {
    (void)verbose;
    // validateWrapperGlobalInputs (wrapper:1025-1075): validation failures leave doseData untouched
    if (!ctData) return set_error(PBS_E_INVALID, "ctData must not be null");
    if (!doseData) return set_error(PBS_E_INVALID, "doseData must not be null");
    if (!beams) return set_error(PBS_E_INVALID, "beamSettings must not be null");
    if (!energy) return set_error(PBS_E_INVALID, "energyData must not be null");
    if (numBeams == 0) return set_error(PBS_E_INVALID, "numBeams must be positive");
    if (ctDims[0] <= 0 || ctDims[1] <= 0 || ctDims[2] <= 0) return set_error(PBS_E_INVALID, "ctDims must be positive");
    if (doseDims[0] <= 0 || doseDims[1] <= 0 || doseDims[2] <= 0) return set_error(PBS_E_INVALID, "doseDims must be positive");
    if (!(ctResolution[0] > 0 && ctResolution[1] > 0 && ctResolution[2] > 0))
        return set_error(PBS_E_INVALID, "ctResolution must be positive");
    if (!(doseResolution[0] > 0 && doseResolution[1] > 0 && doseResolution[2] > 0))
        return set_error(PBS_E_INVALID, "doseResolution must be positive");
    int rc = validate_energy(energy);
    if (rc != PBS_OK) return rc;
    for (size_t b = 0; b < numBeams; ++b) {
        rc = validate_beam(beams[b], b);
        if (rc != PBS_OK) return rc;
    }
    if (nuclearCorrection && !(energy->nucWeightMatrix && energy->nucSqSigmaMatrix))
        return set_error(PBS_E_GEOMETRY, "nuclear_correction=true requested, but the energy struct does not carry nuclear tables");
    if (nuclearCorrection < 0 || nuclearCorrection > 3)
        return set_error(PBS_E_INVALID, "nuclearCorrection must be 0, 1 (GAUSS_FIT), 2 (SOUKUP) or 3 (FLUKA)");
    std::unique_ptr<pbs_engine, void (*)(pbs_engine *)> e(pbs_engine_create(gpuId, energy), pbs_engine_destroy);
    if (!e) return PBS_E_DEVICE;
    // the CT may be a host or a device buffer (texture_ultra_optimized.cu:43-62)
    hipPointerAttribute_t ctAttr{};
    const bool ctOnDevice = hipPointerGetAttributes(&ctAttr, ctData) == hipSuccess && ctAttr.type == hipMemoryTypeDevice;
    (void)hipGetLastError();
    rc = pbs_engine_set_ct(e.get(), ctData, ctDims, ctResolution, ctCorner, ctOnDevice ? 1 : 0);
    if (rc != PBS_OK) return rc;
    rc = pbs_engine_set_dose_grid(e.get(), doseDims, doseResolution, doseCorner);
    if (rc != PBS_OK) return rc;
    rc = pbs_engine_set_nuclear(e.get(), nuclearCorrection);
    if (rc != PBS_OK) return rc;
    const size_t nd = size_t(doseDims[0]) * doseDims[1] * doseDims[2];
    DevBuf<float> dose;
    HIP_TRY(dose.ensure(nd));
    HIP_TRY(hipMemcpy(dose.p, doseData, nd * sizeof(float), hipMemcpyHostToDevice));
    for (size_t b = 0; b < numBeams; ++b) {
        rc = pbs_engine_beam_dose(e.get(), &beams[b], dose.p);
        if (rc != PBS_OK) return rc;
    }
    rc = pbs_engine_sync(e.get());  // reports a halo radius overflow before the dose is copied back
    if (rc != PBS_OK) return rc;
    HIP_TRY(hipMemcpy(doseData, dose.p, nd * sizeof(float), hipMemcpyDeviceToHost));
    return PBS_OK;
}

// ---------------------------------------------------------------------------------------
// cudaFinalDose
// ---------------------------------------------------------------------------------------
int pbs_final_dose_bind(pbs_engine *e, const float *layerEnergies, const float *layerSigmas, int numLayers,
                        float raySpacing) {
    if (!e || !layerEnergies || !layerSigmas || numLayers <= 0)
        return set_error(PBS_E_INVALID, "pbs_final_dose_bind: null argument");
    e->fdEnergies.assign(layerEnergies, layerEnergies + numLayers);
    e->fdSigmas.assign(layerSigmas, layerSigmas + 2 * numLayers);
    e->fdRaySpacing = raySpacing;
    std::lock_guard<std::mutex> lk(g_finalDoseMutex);
    g_finalDoseEngine = e;
    return PBS_OK;
}

int pbs_final_dose_roi(const vec3i *roi, int n, vec3f res, vec3f corner, vec3f bdir, vec3f src, int numLayers,
                       vec3f *minC, vec3f *maxC, vec3i *dims, CPBGrid *cpb) {
    // calculateROIRangeKernel (cuda_final_dose.cu:66-129), margins (5, 5, 2)
    vec3f mn{0.f, 0.f, 0.f}, mx{0.f, 0.f, 0.f};
    vec3i d{0, 0, 0};
    if (n > 0 && roi) {
        vec3i lo = roi[0], hi = roi[0];
        for (int i = 1; i < n; ++i) {
            lo.x = std::min(lo.x, roi[i].x); lo.y = std::min(lo.y, roi[i].y); lo.z = std::min(lo.z, roi[i].z);
            hi.x = std::max(hi.x, roi[i].x); hi.y = std::max(hi.y, roi[i].y); hi.z = std::max(hi.z, roi[i].z);
        }
        mn = vec3f{corner.x + lo.x * res.x, corner.y + lo.y * res.y, corner.z + lo.z * res.z};
        mx = vec3f{corner.x + (hi.x + 1) * res.x, corner.y + (hi.y + 1) * res.y, corner.z + (hi.z + 1) * res.z};
        mn.x -= 5.0f; mn.y -= 5.0f; mn.z -= 2.0f;
        mx.x += 5.0f; mx.y += 5.0f; mx.z += 2.0f;
        d = vec3i{hi.x - lo.x + 1, hi.y - lo.y + 1, hi.z - lo.z + 1};
    }
    if (minC) *minC = mn;
    if (maxC) *maxC = mx;
    if (dims) *dims = d;
    // checkBeamROIIntersectionKernel (cuda_final_dose.cu:132-152)
    const float cx = (mn.x + mx.x) * 0.5f, cy = (mn.y + mx.y) * 0.5f, cz = (mn.z + mx.z) * 0.5f;
    const float dp = bdir.x * (cx - src.x) + bdir.y * (cy - src.y) + bdir.z * (cz - src.z);
    // CPB grid (cuda_final_dose.cu:302-308)
    if (cpb) {
        cpb->corner = vec3f{mn.x, mn.y, 0.0f};
        cpb->resolution = vec3f{0.5f, 0.5f, 1.0f};
        cpb->dims = vec3i{int((mx.x - mn.x) / 0.5f), int((mx.y - mn.y) / 0.5f), numLayers};
        cpb->nLayers = numLayers;
    }
    return dp > 0.0f ? 1 : 0;
}

int cudaFinalDose(const float *subspotData, const vec3i *roiIndices, int numROIIndices, vec3f doseGridResolution,
                  vec3f doseGridCorner, vec3f beamDirection, vec3f sourcePosition, float sad, float refPlaneZ,
                  int numLayers, int maxSubspotsPerLayer, float *finalDose, vec3i doseGridDims) {
    (void)refPlaneZ;
    if (!subspotData || !finalDose || numLayers <= 0 || maxSubspotsPerLayer <= 0)
        return set_error(PBS_E_INVALID, "cudaFinalDose: null or empty argument");
    std::vector<vec3i> roi(std::max(numROIIndices, 0));
    if (numROIIndices > 0) {
        if (!roiIndices) return set_error(PBS_E_INVALID, "cudaFinalDose: roiIndices must not be null");
        hipPointerAttribute_t at{};
        const bool dev = hipPointerGetAttributes(&at, roiIndices) == hipSuccess && at.type == hipMemoryTypeDevice;
        (void)hipGetLastError();
        HIP_TRY(hipMemcpy(roi.data(), roiIndices, sizeof(vec3i) * numROIIndices,
                          dev ? hipMemcpyDeviceToHost : hipMemcpyHostToHost));
    }
    vec3f mn, mx;
    vec3i rd;
    CPBGrid cpb;
    const int hit = pbs_final_dose_roi(roi.data(), numROIIndices, doseGridResolution, doseGridCorner, beamDirection,
                                       sourcePosition, numLayers, &mn, &mx, &rd, &cpb);
    if (!hit) return PBS_OK;  // the reference returns before touching finalDose
    std::lock_guard<std::mutex> lk(g_finalDoseMutex);
    pbs_engine *e = g_finalDoseEngine;
    if (!e || int(e->fdEnergies.size()) != numLayers)
        return set_error(PBS_E_INVALID, "cudaFinalDose: bind CT, LUTs and layer energies with pbs_final_dose_bind");
    const size_t nd = size_t(doseGridDims.x) * doseGridDims.y * doseGridDims.z;
    HIP_TRY(hipMemset(finalDose, 0, nd * sizeof(float)));
    std::vector<float> sub(size_t(numLayers) * maxSubspotsPerLayer * 5);
    {
        hipPointerAttribute_t at{};
        const bool dev = hipPointerGetAttributes(&at, subspotData) == hipSuccess && at.type == hipMemoryTypeDevice;
        (void)hipGetLastError();
        HIP_TRY(hipMemcpy(sub.data(), subspotData, sub.size() * sizeof(float),
                          dev ? hipMemcpyDeviceToHost : hipMemcpyHostToHost));
    }
    const int dd[3] = {doseGridDims.x, doseGridDims.y, doseGridDims.z};
    const float dr[3] = {doseGridResolution.x, doseGridResolution.y, doseGridResolution.z};
    const float dc[3] = {doseGridCorner.x, doseGridCorner.y, doseGridCorner.z};
    int rc = pbs_engine_set_dose_grid(e, dd, dr, dc);
    if (rc != PBS_OK) return rc;
    pbs_beam_settings b{};
    b.numLayers = numLayers;
    b.energies = e->fdEnergies.data();
    b.spotSigmas = e->fdSigmas.data();
    b.raySpacing[0] = b.raySpacing[1] = cpb.resolution.x;
    b.maxSubspotsPerLayer = maxSubspotsPerLayer;
    b.subspotData = sub.data();
    const float bdn = std::sqrt(beamDirection.x * beamDirection.x + beamDirection.y * beamDirection.y +
                                beamDirection.z * beamDirection.z);
    b.beamDirection[0] = beamDirection.x; b.beamDirection[1] = beamDirection.y; b.beamDirection[2] = beamDirection.z;
    b.bmXDirection[0] = 1.0f; b.bmYDirection[1] = 1.0f;  // the reference's fixed lateral axes
    b.sourcePosition[0] = sourcePosition.x; b.sourcePosition[1] = sourcePosition.y; b.sourcePosition[2] = sourcePosition.z;
    b.sad = sad;
    if (!(bdn > 0.0f)) return set_error(PBS_E_INVALID, "cudaFinalDose: zero beam direction");
    // The ROI CPB grid is in world x/y; the gantry frame has y = -bmY, so its y range flips.
    CpbOverride ov{};
    ov.corner[0] = cpb.corner.x;
    ov.corner[1] = -(cpb.corner.y + float(cpb.dims.y) * cpb.resolution.y);
    ov.res[0] = cpb.resolution.x; ov.res[1] = cpb.resolution.y;
    ov.dims[0] = cpb.dims.x; ov.dims[1] = cpb.dims.y;
    // subspot positions of the world-frame CPB convolution, expressed in that gantry frame
    const float geom[13] = {0.f, 0.f, -1.f, 1.f, 0.f, 0.f, 0.f, -1.f, 0.f, 0.f, 0.f, sad, sad};
    rc = engine_beam_dose(e, &b, finalDose, &ov, geom);
    if (rc != PBS_OK) return rc;
    return pbs_engine_sync(e);
}

// ---------------------------------------------------------------------------------------
// batched / served forms
// ---------------------------------------------------------------------------------------
}  // extern "C"

// Deep host copy of an energy table set: the multi-beam batch creates the engines of a table
// set when a batch needs more of them, after the caller's arrays may have gone.
struct HostEnergy {
    pbs_energy_struct e{};
    std::vector<float> energies, peaks, scales, cidd, dens, sp, rrl, nucW, nucS2;
    explicit HostEnergy(const pbs_energy_struct &src) : e(src) {
        const size_t nC = size_t(src.nEnergies) * size_t(src.nEnergySamples);
        auto cp = [](std::vector<float> &v, const float *q, size_t n) -> const float * {
            if (!q) return nullptr;
            v.assign(q, q + n);
            return v.data();
        };
        e.energiesPerU = cp(energies, src.energiesPerU, size_t(src.nEnergies));
        e.peakDepths = cp(peaks, src.peakDepths, size_t(src.nEnergies));
        e.scaleFacts = cp(scales, src.scaleFacts, size_t(src.nEnergies));
        e.ciddMatrix = cp(cidd, src.ciddMatrix, nC);
        e.densityVector = cp(dens, src.densityVector, size_t(src.nDensitySamples));
        e.spVector = cp(sp, src.spVector, size_t(src.nSpSamples));
        e.rRlVector = cp(rrl, src.rRlVector, size_t(src.nRRlSamples));
        e.nucWeightMatrix = cp(nucW, src.nucWeightMatrix, nC);
        e.nucSqSigmaMatrix = cp(nucS2, src.nucSqSigmaMatrix, nC);
    }
};

struct pbs_batch {
    int device = 0;
    std::vector<pbs_engine *> lanes;
    pbs_engine *owner = nullptr;  // holds the CT every lane borrows
    DevBuf<float> laneDose;  // [lanes][dose voxels] per-lane accumulators of the lane form
    size_t doseN = 0;
    size_t next = 0;               // round-robin cursor of the served form
    long long tickets = 0;
    // a submitted request: its ticket, the event after its work, its dose buffer and its lane
    struct Pending {
        long long ticket;
        hipEvent_t done;
        const float *dose;
        int lane;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> freeEvents;
    hipEvent_t laneDone[16] = {};
    // multi-beam form: per energy-table set its host copy, nuclear model and the engines that
    // hold one beam's buffers each; the batch stream and the launch table of its passes
    std::vector<std::unique_ptr<HostEnergy>> tables;
    std::vector<int> tableNuclear;
    std::vector<std::vector<pbs_engine *>> slots;
    hipStream_t stream = nullptr;
    DevBuf<unsigned char> devTable;
    unsigned char *hostTable = nullptr;  // pinned
    size_t hostTableBytes = 0;
    hipEvent_t tableDone = nullptr;  // the launch table's copy of the last multi-beam call
    bool tablePending = false;
    bool multiUsed = false;
    int lastLaunches = 0;
    // the passes' HIP graph and the launch list and table it was captured with
    hipGraphExec_t graphExec = nullptr;
    std::vector<RecLaunch> graphLaunches;
    const unsigned char *graphTable = nullptr;
    int graphCaptures = 0;
    ForkJoin fj{};  // side streams of the independent launches (radius classes)
    bool haveCt = false;
    int ctDims[3] = {};
    float ctRes[3] = {}, ctCorner[3] = {};
    bool haveDose = false;
    int doseDims[3] = {};
    float doseRes[3] = {}, doseCorner[3] = {};
};

namespace {

__global__ void k_add_lanes(float *__restrict__ dose, const float *__restrict__ lanes, int numLanes, size_t n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
        float s = 0.0f;
        for (int l = 0; l < numLanes; ++l) s += lanes[size_t(l) * n + i];
        dose[i] += s;
    }
}

__global__ void k_zero(float *__restrict__ v, size_t n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        v[i] = 0.0f;
}

}  // namespace

extern "C" {

pbs_batch *pbs_batch_create(int gpuId, const pbs_energy_struct *energy, int lanes) {
    if (lanes < 1 || lanes > 16) {
        set_error(PBS_E_INVALID, "lanes must be 1..16");
        return nullptr;
    }
    std::unique_ptr<pbs_batch> b(new pbs_batch());
    b->device = gpuId;
    b->owner = engine_create(gpuId, energy, true);
    if (!b->owner) return nullptr;
    b->tables.emplace_back(new HostEnergy(*energy));
    b->tableNuclear.push_back(0);
    b->slots.emplace_back();
    bool ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&b->tableDone, kOrderEventFlags) == hipSuccess &&
              hipEventCreateWithFlags(&b->fj.fork, kOrderEventFlags) == hipSuccess;
    for (int i = 0; ok && i < kSideStreams; ++i)
        ok = hipStreamCreateWithFlags(&b->fj.side[i], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&b->fj.join[i], kOrderEventFlags) == hipSuccess;
    if (!ok) {
        set_error(PBS_E_DEVICE, "hipStreamCreate failed");
        pbs_engine_destroy(b->owner);
        return nullptr;  // the streams and events made so far stay with the process
    }
    for (int i = 0; i < lanes; ++i) {
        pbs_engine *e = engine_create(gpuId, energy, true);
        if (!e) {
            for (pbs_engine *x : b->lanes) pbs_engine_destroy(x);
            pbs_engine_destroy(b->owner);
            return nullptr;
        }
        b->lanes.push_back(e);
        if (hipEventCreateWithFlags(&b->laneDone[i], kOrderEventFlags) != hipSuccess) {
            set_error(PBS_E_DEVICE, "hipEventCreate failed");
            return nullptr;
        }
    }
    return b.release();
}

void pbs_batch_destroy(pbs_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (pbs_engine *e : b->lanes) pbs_engine_destroy(e);
    for (auto &v : b->slots)
        for (pbs_engine *e : v) pbs_engine_destroy(e);
    pbs_engine_destroy(b->owner);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->tableDone) (void)hipEventDestroy(b->tableDone);
    if (b->graphExec) (void)hipGraphExecDestroy(b->graphExec);
    for (int i = 0; i < kSideStreams; ++i) {
        if (b->fj.side[i]) (void)hipStreamDestroy(b->fj.side[i]);
        if (b->fj.join[i]) (void)hipEventDestroy(b->fj.join[i]);
    }
    if (b->fj.fork) (void)hipEventDestroy(b->fj.fork);
    if (b->hostTable) (void)hipHostFree(b->hostTable);
    for (auto &p : b->pending) (void)hipEventDestroy(p.done);
    for (hipEvent_t ev : b->freeEvents) (void)hipEventDestroy(ev);
    for (int i = 0; i < 16; ++i)
        if (b->laneDone[i]) (void)hipEventDestroy(b->laneDone[i]);
    delete b;
}

int pbs_batch_lanes(pbs_batch *b) { return b ? int(b->lanes.size()) : 0; }

int pbs_batch_set_ct(pbs_batch *b, const float *ct, const int dims[3], const float res[3], const float corner[3],
                     int ctIsDevice) {
    if (!b) return set_error(PBS_E_INVALID, "null batch");
    int rc = pbs_engine_set_ct(b->owner, ct, dims, res, corner, ctIsDevice);
    if (rc != PBS_OK) return rc;
    for (pbs_engine *e : b->lanes) {
        rc = pbs_engine_set_ct(e, b->owner->ctPtr, dims, res, corner, 2);
        if (rc != PBS_OK) return rc;
    }
    for (auto &v : b->slots)
        for (pbs_engine *e : v) {
            rc = pbs_engine_set_ct(e, b->owner->ctPtr, dims, res, corner, 2);
            if (rc != PBS_OK) return rc;
        }
    b->haveCt = true;
    for (int i = 0; i < 3; ++i) { b->ctDims[i] = dims[i]; b->ctRes[i] = res[i]; b->ctCorner[i] = corner[i]; }
    return PBS_OK;
}

int pbs_batch_set_dose_grid(pbs_batch *b, const int dims[3], const float res[3], const float corner[3]) {
    if (!b) return set_error(PBS_E_INVALID, "null batch");
    for (pbs_engine *e : b->lanes) {
        const int rc = pbs_engine_set_dose_grid(e, dims, res, corner);
        if (rc != PBS_OK) return rc;
    }
    for (auto &v : b->slots)
        for (pbs_engine *e : v) {
            const int rc = pbs_engine_set_dose_grid(e, dims, res, corner);
            if (rc != PBS_OK) return rc;
        }
    b->doseN = size_t(dims[0]) * dims[1] * dims[2];
    b->haveDose = true;
    for (int i = 0; i < 3; ++i) { b->doseDims[i] = dims[i]; b->doseRes[i] = res[i]; b->doseCorner[i] = corner[i]; }
    return PBS_OK;
}

int pbs_batch_set_nuclear(pbs_batch *b, int model) {
    if (!b) return set_error(PBS_E_INVALID, "null batch");
    for (pbs_engine *e : b->lanes) {
        const int rc = pbs_engine_set_nuclear(e, model);
        if (rc != PBS_OK) return rc;
    }
    for (pbs_engine *e : b->slots[0]) {
        const int rc = pbs_engine_set_nuclear(e, model);
        if (rc != PBS_OK) return rc;
    }
    b->tableNuclear[0] = model;
    return PBS_OK;
}

static int batch_dose_lanes(pbs_batch *b, const pbs_beam_settings *beams, size_t numBeams, float *doseDev) {
    if (b->doseN == 0) return set_error(PBS_E_INVALID, "no dose grid bound to the batch");
    HIP_TRY(hipSetDevice(b->device));
    const int L = int(b->lanes.size());
    const size_t n = b->doseN;
    HIP_TRY(b->laneDose.ensure(n * L));
    float *acc = b->laneDose.p;
    hipStream_t s0 = b->lanes[0]->stream;
    // lanes start from zero once every lane's earlier work on the buffer has finished
    for (int l = 0; l < L; ++l) {
        HIP_TRY(hipEventRecord(b->laneDone[l], b->lanes[l]->stream));
        HIP_TRY(hipStreamWaitEvent(s0, b->laneDone[l], 0));
    }
    hipLaunchKernelGGL(k_zero, dim3(2048), dim3(256), 0, s0, acc, n * L);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(b->laneDone[0], s0));
    for (int l = 1; l < L; ++l) HIP_TRY(hipStreamWaitEvent(b->lanes[l]->stream, b->laneDone[0], 0));
    for (size_t i = 0; i < numBeams; ++i) {
        const int l = int(i % size_t(L));
        const int rc = pbs_engine_beam_dose(b->lanes[l], &beams[i], acc + size_t(l) * n);
        if (rc != PBS_OK) return rc;
    }
    // join the lanes on lane 0 and add them in lane order
    for (int l = 1; l < L; ++l) {
        HIP_TRY(hipEventRecord(b->laneDone[l], b->lanes[l]->stream));
        HIP_TRY(hipStreamWaitEvent(s0, b->laneDone[l], 0));
    }
    hipLaunchKernelGGL(k_add_lanes, dim3(2048), dim3(256), 0, s0, doseDev, acc, L, n);
    HIP_TRY(hipGetLastError());
    return PBS_OK;
}

// Engine j of table set t for a multi-beam call (created on first use and bound to the batch's
// CT, dose grid and the table's nuclear model).
static pbs_engine *batch_slot(pbs_batch *b, int t, size_t j) {
    std::vector<pbs_engine *> &v = b->slots[size_t(t)];
    while (v.size() <= j) {
        pbs_engine *e = engine_create(b->device, &b->tables[size_t(t)]->e, true);
        if (!e) return nullptr;
        v.push_back(e);
        int rc = PBS_OK;
        if (b->haveCt) rc = pbs_engine_set_ct(e, b->owner->ctPtr, b->ctDims, b->ctRes, b->ctCorner, 2);
        if (rc == PBS_OK && b->haveDose) rc = pbs_engine_set_dose_grid(e, b->doseDims, b->doseRes, b->doseCorner);
        if (rc == PBS_OK) rc = pbs_engine_set_nuclear(e, b->tableNuclear[size_t(t)]);
        if (rc != PBS_OK) return nullptr;
    }
    return v[j];
}

// The multi-beam form: every beam's chain is recorded on its own engine (planning, buffers and
// staging as for one beam, no launch), then the batch stream issues each pass once for all
// beams (stage_records, issue_records): the launches of a batch do not grow with its beam count.
static int batch_dose_multi(pbs_batch *b, const pbs_beam_settings *beams, const int *tableIds, size_t numBeams,
                            float *doseDev) {
    if (!b->haveCt) return set_error(PBS_E_INVALID, "no CT bound to the batch");
    HIP_TRY(hipSetDevice(b->device));
    Recorder rec;
    std::vector<size_t> used(b->tables.size(), 0);
    std::vector<pbs_engine *> engines;
    for (size_t i = 0; i < numBeams; ++i) {
        const int t = tableIds ? tableIds[i] : 0;
        if (t < 0 || size_t(t) >= b->tables.size())
            return set_error(PBS_E_INVALID, "beam " + std::to_string(i) + ": unknown table set " + std::to_string(t));
        pbs_engine *e = batch_slot(b, t, used[size_t(t)]++);
        if (!e) return PBS_E_DEVICE;
        engines.push_back(e);
        rec.beam = int(i);
        rec.seq = 0;
        rec.set_stage(0);
        t_rec = &rec;
        const int rc = engine_beam_dose(e, &beams[i], doseDev, nullptr, nullptr);
        t_rec = nullptr;
        if (rc != PBS_OK) return rc;
    }
    const size_t need = records_table_bytes(rec.recs);
    if (b->tablePending) {
        HIP_TRY(hipEventSynchronize(b->tableDone));
        b->tablePending = false;
    }
    if (need > b->hostTableBytes) {
        // the device table of earlier calls may still be read: grow only on an idle stream
        HIP_TRY(hipStreamSynchronize(b->stream));
        if (b->hostTable) (void)hipHostFree(b->hostTable);
        b->hostTable = nullptr;
        b->hostTableBytes = 0;
        const size_t cap = std::max(need + need / 2, size_t(1) << 16);
        HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&b->hostTable), cap));
        b->hostTableBytes = cap;
        HIP_TRY(b->devTable.ensure(cap));
    }
    std::vector<RecLaunch> launches;
    const size_t bytes = stage_records(rec.recs, b->hostTable, launches);
    if (bytes) HIP_TRY(hipMemcpyAsync(b->devTable.p, b->hostTable, bytes, hipMemcpyHostToDevice, b->stream));
    // the passes as one HIP graph: captured when the launch list is new, replayed while it repeats
    // (the per-beam arguments live in the table copied above, so a replay serves new beams of the
    // same shape)
    if (!(b->graphExec && launches == b->graphLaunches && b->devTable.p == b->graphTable)) {
        if (b->graphExec) (void)hipGraphExecDestroy(b->graphExec);
        b->graphExec = nullptr;
        hipGraph_t graph = nullptr;
        HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const hipError_t ie = issue_records(launches, b->devTable.p, b->stream, &b->fj);
        const hipError_t ce = hipStreamEndCapture(b->stream, &graph);
        HIP_TRY(ie);
        HIP_TRY(ce);
        const hipError_t xe = hipGraphInstantiate(&b->graphExec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIP_TRY(xe);
        b->graphLaunches = launches;
        b->graphTable = b->devTable.p;
        ++b->graphCaptures;
    }
    HIP_TRY(hipGraphLaunch(b->graphExec, b->stream));
    HIP_TRY(hipEventRecord(b->tableDone, b->stream));
    b->tablePending = true;
    // each engine's staging is read by this batch's upload pass: its next beam waits for it
    for (pbs_engine *e : engines) {
        HIP_TRY(hipEventRecord(e->uploadDone, b->stream));
        e->uploadPending = true;
    }
    b->lastLaunches = int(launches.size());
    b->multiUsed = true;
    return PBS_OK;
}

int pbs_batch_dose(pbs_batch *b, const pbs_beam_settings *beams, size_t numBeams, float *doseDev) {
    if (!b || !beams || !doseDev) return set_error(PBS_E_INVALID, "null argument");
    if (b->doseN == 0) return set_error(PBS_E_INVALID, "no dose grid bound to the batch");
    b->lastLaunches = 0;
    const int rc = batch_dose_multi(b, beams, nullptr, numBeams, doseDev);
    if (rc != PBS_E_UNSUPPORTED) return rc;
    return batch_dose_lanes(b, beams, numBeams, doseDev);  // beams the multi-beam passes do not cover
}

int pbs_batch_add_table(pbs_batch *b, const pbs_energy_struct *energy, int nuclearModel) {
    if (!b || !energy) return set_error(PBS_E_INVALID, "null argument");
    const int rc = validate_energy(energy);
    if (rc != PBS_OK) return rc;
    if (nuclearModel < 0 || nuclearModel > 3)
        return set_error(PBS_E_INVALID, "nuclearModel must be 0 (off), 1 (GAUSS_FIT), 2 (SOUKUP) or 3 (FLUKA)");
    if (nuclearModel && !(energy->nucWeightMatrix && energy->nucSqSigmaMatrix))
        return set_error(PBS_E_GEOMETRY, "a nuclear model needs the table set's nuclear tables");
    b->tables.emplace_back(new HostEnergy(*energy));
    b->tableNuclear.push_back(nuclearModel);
    b->slots.emplace_back();
    return int(b->tables.size()) - 1;
}

int pbs_batch_dose_tables(pbs_batch *b, const pbs_beam_settings *beams, const int *tableIds, size_t numBeams,
                          float *doseDev) {
    if (!b || !beams || !doseDev) return set_error(PBS_E_INVALID, "null argument");
    if (b->doseN == 0) return set_error(PBS_E_INVALID, "no dose grid bound to the batch");
    b->lastLaunches = 0;
    return batch_dose_multi(b, beams, tableIds, numBeams, doseDev);
}

int pbs_batch_last_launches(pbs_batch *b) { return b ? b->lastLaunches : 0; }

int pbs_batch_graph_captures(pbs_batch *b) { return b ? b->graphCaptures : 0; }

int pbs_batch_submit(pbs_batch *b, const pbs_beam_settings *beam, float *doseDev, long long *ticket) {
    if (!b || !beam || !doseDev) return set_error(PBS_E_INVALID, "null argument");
    HIP_TRY(hipSetDevice(b->device));
    // The transfer adds into the dose without atomics, so two requests into one buffer must not
    // run on two lanes at once: a buffer with a request still queued goes to that request's lane,
    // whose stream orders the two; other buffers take the lanes in turn.
    int l = -1;
    for (auto it = b->pending.rbegin(); it != b->pending.rend(); ++it) {
        if (it->dose != doseDev) continue;
        const hipError_t q = hipEventQuery(it->done);
        if (q == hipErrorNotReady) l = it->lane;
        else HIP_TRY(q);
        break;
    }
    if (l < 0) l = int(b->next++ % b->lanes.size());
    pbs_engine *e = b->lanes[l];
    const int rc = pbs_engine_beam_dose(e, beam, doseDev);
    if (rc != PBS_OK) return rc;
    hipEvent_t ev;
    if (!b->freeEvents.empty()) {
        ev = b->freeEvents.back();
        b->freeEvents.pop_back();
    } else {
        HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    }
    HIP_TRY(hipEventRecord(ev, e->stream));
    const long long t = ++b->tickets;
    b->pending.push_back({t, ev, doseDev, l});
    if (ticket) *ticket = t;
    return PBS_OK;
}

int pbs_batch_query(pbs_batch *b, long long ticket) {
    if (!b) return set_error(PBS_E_INVALID, "null batch");
    for (size_t i = 0; i < b->pending.size(); ++i) {
        if (b->pending[i].ticket != ticket) continue;
        const hipError_t q = hipEventQuery(b->pending[i].done);
        if (q == hipErrorNotReady) return 0;
        HIP_TRY(q);
        b->freeEvents.push_back(b->pending[i].done);
        b->pending.erase(b->pending.begin() + long(i));
        return 1;
    }
    return 1;  // unknown or already collected: done
}

int pbs_batch_wait(pbs_batch *b, long long ticket) {
    if (!b) return set_error(PBS_E_INVALID, "null batch");
    for (size_t i = 0; i < b->pending.size(); ++i) {
        if (b->pending[i].ticket != ticket) continue;
        HIP_TRY(hipEventSynchronize(b->pending[i].done));
        b->freeEvents.push_back(b->pending[i].done);
        b->pending.erase(b->pending.begin() + long(i));
        break;
    }
    return PBS_OK;
}

int pbs_batch_sync(pbs_batch *b) {
    if (!b) return set_error(PBS_E_INVALID, "null batch");
    int rc = PBS_OK;
    for (pbs_engine *e : b->lanes) {
        const int r = pbs_engine_sync(e);
        if (r != PBS_OK && rc == PBS_OK) rc = r;
    }
    if (b->multiUsed) {
        // the multi-beam passes ran on the batch stream; then each engine's halo check
        HIP_TRY(hipSetDevice(b->device));
        HIP_TRY(hipStreamSynchronize(b->stream));
        b->tablePending = false;
        for (auto &v : b->slots)
            for (pbs_engine *e : v) {
                const int r = pbs_engine_sync(e);
                if (r != PBS_OK && rc == PBS_OK) rc = r;
            }
    }
    for (auto &p : b->pending) b->freeEvents.push_back(p.done);
    b->pending.clear();
    return rc;
}

}  // extern "C\"

// ---------------------------------------------------------------------------------------
// served form: a request server with resident CTs and beams, admission by device-memory
// footprint and continuous batching over engine slots
// ---------------------------------------------------------------------------------------
namespace {

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// Deep copy of a pbs_beam_settings (the caller's arrays may go away after submit).
struct BeamCopy {
    pbs_beam_settings b{};
    std::vector<float> energies, sigmas, subspots, positions, weights, dirs, cutoffs;
    std::vector<int> counts, roi;
    void set(const pbs_beam_settings &in) {
        b = in;
        const int L = in.numLayers;
        auto cp = [](std::vector<float> &v, const float *p, size_t n) -> const float * {
            if (!p) return nullptr;
            v.assign(p, p + n);
            return v.data();
        };
        b.energies = cp(energies, in.energies, size_t(L));
        b.spotSigmas = cp(sigmas, in.spotSigmas, size_t(L) * 2);
        b.subspotData = cp(subspots, in.subspotData, size_t(L) * std::max(in.maxSubspotsPerLayer, 0) * 5);
        b.spotPositions = cp(positions, in.spotPositions, size_t(std::max(in.numSpots, 0)) * 2);
        b.spotWeights = cp(weights, in.spotWeights, size_t(std::max(in.numSpots, 0)));
        b.spotBeamDirections = cp(dirs, in.spotBeamDirections, size_t(std::max(in.numSpots, 0)) * 3);
        b.layerLongitudinalCutoffs = cp(cutoffs, in.layerLongitudinalCutoffs, size_t(L));
        if (in.layerSpotCounts) {
            counts.assign(in.layerSpotCounts, in.layerSpotCounts + L);
            b.layerSpotCounts = counts.data();
        }
        if (in.roiLinearIndices && in.numRoiLinearIndices > 0) {
            roi.assign(in.roiLinearIndices, in.roiLinearIndices + in.numRoiLinearIndices);
            b.roiLinearIndices = roi.data();
        }
    }
};

struct ServerCt {
    DevBuf<float> ct;  // HU+1000, resident
    int dims[3] = {0, 0, 0};
    float res[3] = {0, 0, 0}, corner[3] = {0, 0, 0};
    int users = 0;     // queued or running requests and resident beams on this CT
};

struct ServerReq {
    long long ticket = 0;
    int kind = 0;  // 0 full beam, 1 weight-only update of a resident beam
    int ctId = -1;
    int doseDims[3] = {0, 0, 0};
    float doseRes[3] = {0, 0, 0}, doseCorner[3] = {0, 0, 0};
    BeamCopy beam;
    bool keep = false;                 // full beam whose state stays resident (handle = ticket)
    int numSpots = 0;
    long long stateTicket = 0;         // weight-only: the resident beam
    std::vector<float> weights;
    int status = 0;                    // 0 queued, 1 running, 2 done, < 0 error code
    std::string error;
    int slot = -1;
    float *dose = nullptr;
    size_t doseN = 0;
    size_t doseCharge = 0, stateCharge = 0;  // budget bytes held for the dose buffer / resident state
    hipEvent_t done = nullptr;
    double tSubmit = 0, tStart = 0, tDone = 0;
};

}  // namespace

struct pbs_server {
    int device = 0;
    std::vector<pbs_engine *> slots;
    std::vector<long long> slotTicket;   // running request per slot, 0 = free
    std::vector<size_t> slotWork;        // device bytes each slot's engine holds for its work buffers
    std::map<int, std::unique_ptr<ServerCt>> cts;
    int nextCt = 1;
    std::map<long long, std::unique_ptr<BeamState>> states;  // resident beams by their first ticket
    std::map<long long, int> stateCt;
    std::map<long long, std::unique_ptr<ServerReq>> reqs;
    std::deque<long long> queue;
    long long tickets = 0;
    size_t budget = 0, used = 0;  // admission budget and bytes charged (work buffers, doses, states)
    hipStream_t copyStream = nullptr;
    std::vector<hipEvent_t> freeEvents;
    // dose buffers of fetched results, by size, for the next requests: a buffer returns here only
    // after its result copy has completed on the host, so reuse needs no stream-ordered
    // allocator; their bytes stay charged to the budget until freed
    std::multimap<size_t, float *> idleDose;
    size_t idleBytes = 0;
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv, doneCv;
    bool stop = false;
};

namespace {

size_t dose_bytes(const int d[3]) { return size_t(d[0]) * d[1] * d[2] * sizeof(float); }

// Device bytes an engine's work buffers need for one beam (the GridWork and per-beam arrays of
// engine_beam_dose), from the host plan.
size_t work_bytes(const BeamPlan &p, int nSub) {
    const size_t L = size_t(p.numLayers), K = size_t(p.steps);
    const size_t nray = size_t(p.rayDims[0]) * p.rayDims[1], nsup = size_t(p.superpDims[0]) * p.superpDims[1];
    const size_t nt = nsup / 256, maxItems = K * nt;
    const size_t tailCap = std::min(L, K) * std::max<size_t>(1, (K + 1) / 2);
    const size_t bev = size_t(p.superpDims[0] + 64) * (p.superpDims[1] + 64);
    size_t b = 0;
    b += (2 * nray * K + nray * L * 2 + K) * 4;                    // density, wepl, rayW, firstPassive, weplMin
    b += nsup * K * L * 8;                                         // iddSigma
    b += L * K * nt * 4 * 2;                                       // tile codes, slot layer lists
    b += (maxItems + tailCap * nt) * kSlotFloats * 4;              // item blocks
    b += (K + tailCap) * bev * 4;                                  // BEV planes
    b += size_t(kNumClasses) * (maxItems + tailCap * nt) * 8;      // work-unit lists
    b += size_t(p.spotDims[0]) * p.spotDims[1] * L * 4 * (1 + size_t(nSub));  // lattice, convolution
    return b;
}

size_t state_bytes(const BeamPlan &p, int maxSub) {
    const size_t L = size_t(p.numLayers), K = size_t(p.steps);
    const size_t nray = size_t(p.rayDims[0]) * p.rayDims[1], nsup = size_t(p.superpDims[0]) * p.superpDims[1];
    return nsup * K * L * 8 + nray * L * 4 + K * 4 + L * size_t(std::max(maxSub, 0)) * 20 + sizeof(BeamScalars);
}

void finish_error(pbs_server *sv, ServerReq &r, int code, const std::string &msg) {
    if (r.kind == 0) sv->cts[r.ctId]->users--;
    r.status = code;
    r.error = msg;
    r.tDone = now_ms();
    sv->doneCv.notify_all();
}

// Admits the next queued request that fits the budget onto a free slot (called with mu held;
// releases it while the beam is planned and queued). Returns true if one was started.
bool admit_one(pbs_server *sv, std::unique_lock<std::mutex> &lk) {
    int slot = -1;
    for (size_t i = 0; i < sv->slots.size(); ++i)
        if (sv->slotTicket[i] == 0) { slot = int(i); break; }
    if (slot < 0) return false;
    for (size_t qi = 0; qi < sv->queue.size(); ++qi) {
        ServerReq &r = *sv->reqs[sv->queue[qi]];
        // footprint: the dose buffer, the resident state of a kept beam, and the growth of the
        // slot's work buffers
        BeamPlan plan;
        int nSub = 1, maxSub = 0;
        const BeamState *st = nullptr;
        if (r.kind == 1) {
            // an update reads the resident field its beam's first run writes: it starts after that run
            auto base = sv->reqs.find(r.stateTicket);
            if (base != sv->reqs.end() && base->second->status >= 0 && base->second->status != 2) continue;
            auto it = sv->states.find(r.stateTicket);
            if (it == sv->states.end()) {
                sv->queue.erase(sv->queue.begin() + long(qi));
                finish_error(sv, r, PBS_E_INVALID, "unknown or released resident beam");
                return true;
            }
            st = it->second.get();
            plan = st->plan;
            nSub = st->nSub;
            maxSub = st->maxSub;
        } else {
            ServerCt &ct = *sv->cts[r.ctId];
            if (!plan_beam(ct.dims, ct.res, ct.corner, r.doseDims, r.doseRes, r.doseCorner, r.beam.b,
                           sv->slots[size_t(slot)]->energy, plan, nullptr, &sv->slots[size_t(slot)]->env)) {
                sv->queue.erase(sv->queue.begin() + long(qi));
                finish_error(sv, r, PBS_E_GEOMETRY, plan.error);
                return true;
            }
            maxSub = r.beam.b.maxSubspotsPerLayer;
            nSub = (plan.subspotConv && !plan.legacyPath) ? maxSub : 1;
        }
        const size_t need = work_bytes(plan, nSub);
        const size_t grow = need > sv->slotWork[size_t(slot)] ? need - sv->slotWork[size_t(slot)] : 0;
        const size_t doseB = dose_bytes(r.doseDims), stateB = r.keep ? state_bytes(plan, maxSub) : 0;
        auto idle = sv->idleDose.find(doseB);
        size_t fp = (idle != sv->idleDose.end() ? 0 : doseB) + stateB + grow;
        // idle dose buffers of other sizes give their memory back before the request waits for it
        for (auto it = sv->idleDose.begin(); sv->used + fp > sv->budget && it != sv->idleDose.end();) {
            if (it == idle) {
                ++it;
                continue;
            }
            (void)hipFree(it->second);
            sv->used -= std::min(it->first, sv->used);
            sv->idleBytes -= it->first;
            it = sv->idleDose.erase(it);
        }
        if (sv->used + fp > sv->budget) {
            bool anyRunning = false;
            for (long long t : sv->slotTicket) anyRunning |= t != 0;
            if (anyRunning) continue;  // wait for memory; a smaller request further back may fit
            sv->queue.erase(sv->queue.begin() + long(qi));
            finish_error(sv, r, PBS_E_NOMEM, "request exceeds the server's device-memory budget");
            return true;
        }
        sv->used += fp;
        if (idle != sv->idleDose.end()) {
            r.dose = idle->second;
            sv->idleBytes -= idle->first;
            sv->idleDose.erase(idle);
        }
        sv->slotWork[size_t(slot)] += grow;  // work buffers stay with the slot
        r.doseCharge = doseB;
        r.stateCharge = stateB;
        r.slot = slot;
        r.status = 1;
        r.tStart = now_ms();
        sv->slotTicket[size_t(slot)] = r.ticket;
        sv->queue.erase(sv->queue.begin() + long(qi));
        ServerReq *rp = &r;
        std::unique_ptr<BeamState> keepState;
        if (r.keep) keepState.reset(new BeamState());
        lk.unlock();
        // plan and queue the beam on the slot's engine
        pbs_engine *e = sv->slots[size_t(slot)];
        int rc = PBS_OK;
        std::string err;
        hipError_t he = hipSuccess;
        if (!rp->dose) he = hipMalloc(reinterpret_cast<void **>(&rp->dose), dose_bytes(rp->doseDims));
        if (he == hipSuccess) he = hipMemsetAsync(rp->dose, 0, dose_bytes(rp->doseDims), e->stream);
        if (he != hipSuccess) {
            rc = PBS_E_DEVICE;
            err = hipGetErrorString(he);
        }
        rp->doseN = size_t(rp->doseDims[0]) * rp->doseDims[1] * rp->doseDims[2];
        if (rc == PBS_OK) {
            if (rp->kind == 1) {
                (void)pbs_engine_set_dose_grid(e, rp->doseDims, rp->doseRes, rp->doseCorner);
                ServerCt &ct = *sv->cts[rp->ctId];
                (void)pbs_engine_set_ct(e, ct.ct.p, ct.dims, ct.res, ct.corner, 2);
                rc = engine_weights_update(e, *st, rp->weights.data(), int(rp->weights.size()), rp->dose);
            } else {
                ServerCt &ct = *sv->cts[rp->ctId];
                (void)pbs_engine_set_ct(e, ct.ct.p, ct.dims, ct.res, ct.corner, 2);
                (void)pbs_engine_set_dose_grid(e, rp->doseDims, rp->doseRes, rp->doseCorner);
                // the plan made at admission (same CT, dose grid, tables and environment)
                rc = engine_beam_dose(e, &rp->beam.b, rp->dose, nullptr, nullptr, nullptr, nullptr, keepState.get(),
                                      &plan);
            }
            if (rc != PBS_OK) err = g_lastError;
        }
        if (rc == PBS_OK) {
            if (!sv->freeEvents.empty()) {
                rp->done = sv->freeEvents.back();
                sv->freeEvents.pop_back();
            } else if (hipEventCreateWithFlags(&rp->done, hipEventDisableTiming) != hipSuccess) {
                rp->done = nullptr;
            }
            if (!rp->done || hipEventRecord(rp->done, e->stream) != hipSuccess) {
                rc = PBS_E_DEVICE;
                err = "hipEventRecord failed";
            }
        }
        // the engine has planned and queued the beam: the request's host copies are no longer needed
        rp->beam = BeamCopy();
        std::vector<float>().swap(rp->weights);
        lk.lock();
        if (rc != PBS_OK) {
            (void)hipStreamSynchronize(e->stream);
            sv->slotTicket[size_t(slot)] = 0;
            if (rp->dose) {
                sv->idleDose.emplace(rp->doseCharge, rp->dose);  // the stream is idle: reusable at once
                sv->idleBytes += rp->doseCharge;
            } else {
                sv->used -= std::min(rp->doseCharge, sv->used);
            }
            rp->dose = nullptr;
            sv->used -= std::min(rp->stateCharge, sv->used);
            rp->doseCharge = rp->stateCharge = 0;
            finish_error(sv, *rp, rc, err);
        } else if (rp->keep) {
            sv->states[rp->ticket] = std::move(keepState);
            sv->stateCt[rp->ticket] = rp->ctId;
            sv->cts[rp->ctId]->users++;
        }
        return true;
    }
    return false;
}

// Collects finished requests (mu held): frees their slot; the dose stays until fetched.
bool collect(pbs_server *sv) {
    bool any = false;
    for (size_t i = 0; i < sv->slots.size(); ++i) {
        const long long t = sv->slotTicket[i];
        if (!t) continue;
        ServerReq &r = *sv->reqs[t];
        const hipError_t q = hipEventQuery(r.done);
        if (q == hipErrorNotReady) continue;
        r.tDone = now_ms();
        if (q != hipSuccess) {
            r.status = PBS_E_DEVICE;
            r.error = hipGetErrorString(q);
        } else {
            r.status = 2;
        }
        sv->freeEvents.push_back(r.done);
        r.done = nullptr;
        sv->slotTicket[i] = 0;
        if (r.kind == 0) sv->cts[r.ctId]->users--;
        any = true;
    }
    if (any) sv->doneCv.notify_all();
    return any;
}

void worker_loop(pbs_server *sv) {
    (void)hipSetDevice(sv->device);
    std::unique_lock<std::mutex> lk(sv->mu);
    while (!sv->stop) {
        bool progress = collect(sv);
        while (admit_one(sv, lk)) progress = true;
        if (!progress) {
            bool running = false;
            for (long long t : sv->slotTicket) running |= t != 0;
            // poll the running requests' events; sleep until a submit when idle
            if (running)
                sv->cv.wait_for(lk, std::chrono::microseconds(20));
            else
                sv->cv.wait_for(lk, std::chrono::milliseconds(50));
        }
    }
}

int server_request_status(pbs_server *sv, long long ticket, ServerReq **out) {
    auto it = sv->reqs.find(ticket);
    if (it == sv->reqs.end()) return set_error(PBS_E_INVALID, "unknown ticket");
    *out = it->second.get();
    return PBS_OK;
}

}  // namespace

extern "C" {

pbs_server *pbs_server_create(int gpuId, const pbs_energy_struct *energy, int slots, double memoryBudgetBytes) {
    if (slots < 1 || slots > 16) {
        set_error(PBS_E_INVALID, "slots must be 1..16");
        return nullptr;
    }
    if (hipSetDevice(gpuId) != hipSuccess) {
        set_error(PBS_E_DEVICE, "hipSetDevice failed");
        return nullptr;
    }
    std::unique_ptr<pbs_server> sv(new pbs_server());
    sv->device = gpuId;
    for (int i = 0; i < slots; ++i) {
        pbs_engine *e = engine_create(gpuId, energy, true);
        if (!e) {
            for (pbs_engine *x : sv->slots) pbs_engine_destroy(x);
            return nullptr;
        }
        sv->slots.push_back(e);
    }
    sv->slotTicket.assign(size_t(slots), 0);
    sv->slotWork.assign(size_t(slots), 0);
    size_t freeB = 0, totalB = 0;
    (void)hipMemGetInfo(&freeB, &totalB);
    sv->budget = memoryBudgetBytes > 0 ? size_t(memoryBudgetBytes) : size_t(0.8 * double(freeB));
    if (hipStreamCreateWithFlags(&sv->copyStream, hipStreamNonBlocking) != hipSuccess) {
        for (pbs_engine *x : sv->slots) pbs_engine_destroy(x);
        set_error(PBS_E_DEVICE, "hipStreamCreate failed");
        return nullptr;
    }
    pbs_server *raw = sv.release();
    raw->worker = std::thread(worker_loop, raw);
    return raw;
}

void pbs_server_destroy(pbs_server *sv) {
    if (!sv) return;
    {
        std::lock_guard<std::mutex> g(sv->mu);
        sv->stop = true;
    }
    sv->cv.notify_all();
    if (sv->worker.joinable()) sv->worker.join();
    (void)hipSetDevice(sv->device);
    for (pbs_engine *e : sv->slots) (void)hipStreamSynchronize(e->stream);
    for (auto &kv : sv->reqs) {
        if (kv.second->dose) (void)hipFree(kv.second->dose);
        if (kv.second->done) (void)hipEventDestroy(kv.second->done);
    }
    for (hipEvent_t ev : sv->freeEvents) (void)hipEventDestroy(ev);
    for (auto &kv : sv->idleDose) (void)hipFree(kv.second);
    sv->states.clear();
    for (pbs_engine *e : sv->slots) pbs_engine_destroy(e);
    if (sv->copyStream) (void)hipStreamDestroy(sv->copyStream);
    delete sv;
}

int pbs_server_add_ct(pbs_server *sv, const float *ct, const int ctDims[3], const float ctRes[3],
                      const float ctCorner[3], int ctIsDevice, int *ctId) {
    if (!sv || !ct || !ctDims || !ctRes || !ctCorner || !ctId) return set_error(PBS_E_INVALID, "null argument");
    // the CT adapter of the engine API, then a resident copy owned by the server
    std::unique_ptr<ServerCt> c(new ServerCt());
    const size_t n = size_t(ctDims[0]) * ctDims[1] * ctDims[2];
    std::vector<float> host;
    const float *src = ct;
    HIP_TRY(hipSetDevice(sv->device));
    if (ctIsDevice) {
        host.resize(n);
        HIP_TRY(hipMemcpy(host.data(), ct, n * sizeof(float), hipMemcpyDeviceToHost));
        src = host.data();
    }
    std::vector<float> conv;
    if (sv->slots.empty()) return set_error(PBS_E_INVALID, "server has no engine slots");
    int rc = adapt_ct(sv->slots[0], src, n, conv, src);
    if (rc != PBS_OK) return rc;
    HIP_TRY(c->ct.ensure(n));
    HIP_TRY(hipMemcpy(c->ct.p, src, n * sizeof(float), hipMemcpyHostToDevice));
    for (int i = 0; i < 3; ++i) { c->dims[i] = ctDims[i]; c->res[i] = ctRes[i]; c->corner[i] = ctCorner[i]; }
    std::lock_guard<std::mutex> g(sv->mu);
    sv->used += n * sizeof(float);
    *ctId = sv->nextCt++;
    sv->cts[*ctId] = std::move(c);
    return PBS_OK;
}

int pbs_server_remove_ct(pbs_server *sv, int ctId) {
    if (!sv) return set_error(PBS_E_INVALID, "null server");
    std::lock_guard<std::mutex> g(sv->mu);
    auto it = sv->cts.find(ctId);
    if (it == sv->cts.end()) return set_error(PBS_E_INVALID, "unknown CT");
    if (it->second->users > 0) return set_error(PBS_E_INVALID, "CT in use by requests or resident beams");
    sv->used -= it->second->ct.n * sizeof(float);
    sv->cts.erase(it);
    return PBS_OK;
}

int pbs_server_submit(pbs_server *sv, int ctId, const int doseDims[3], const float doseRes[3],
                      const float doseCorner[3], const pbs_beam_settings *beam, int keepState, long long *ticket) {
    if (!sv || !doseDims || !doseRes || !doseCorner || !beam || !ticket) return set_error(PBS_E_INVALID, "null argument");
    int rc = validate_beam(*beam, 0);
    if (rc != PBS_OK) return rc;
    if (doseDims[0] <= 0 || doseDims[1] <= 0 || doseDims[2] <= 0) return set_error(PBS_E_INVALID, "doseDims must be positive");
    std::unique_ptr<ServerReq> r(new ServerReq());
    r->kind = 0;
    r->ctId = ctId;
    for (int i = 0; i < 3; ++i) { r->doseDims[i] = doseDims[i]; r->doseRes[i] = doseRes[i]; r->doseCorner[i] = doseCorner[i]; }
    r->beam.set(*beam);
    r->keep = keepState != 0;
    r->numSpots = beam->numSpots;
    {
        std::lock_guard<std::mutex> g(sv->mu);
        auto it = sv->cts.find(ctId);
        if (it == sv->cts.end()) return set_error(PBS_E_INVALID, "unknown CT");
        it->second->users++;
        r->ticket = ++sv->tickets;
        r->tSubmit = now_ms();
        *ticket = r->ticket;
        sv->queue.push_back(r->ticket);
        sv->reqs[r->ticket] = std::move(r);
    }
    sv->cv.notify_all();
    return PBS_OK;
}

int pbs_server_submit_weights(pbs_server *sv, long long stateTicket, const float *spotWeights, int numSpots,
                              long long *ticket) {
    if (!sv || !spotWeights || !ticket) return set_error(PBS_E_INVALID, "null argument");
    std::unique_ptr<ServerReq> r(new ServerReq());
    r->kind = 1;
    r->stateTicket = stateTicket;
    r->weights.assign(spotWeights, spotWeights + std::max(numSpots, 0));
    {
        std::lock_guard<std::mutex> g(sv->mu);
        // the resident beam may still be queued or running: its first request holds the handle
        auto fr = sv->reqs.find(stateTicket);
        if (fr == sv->reqs.end() || !fr->second->keep || fr->second->status < 0)
            return set_error(PBS_E_INVALID, "unknown or released resident beam");
        if (fr->second->status == 2 && !sv->states.count(stateTicket))
            return set_error(PBS_E_INVALID, "unknown or released resident beam");
        if (numSpots != fr->second->numSpots)
            return set_error(PBS_E_INVALID, "weight count differs from the resident beam's");
        const ServerReq &first = *fr->second;
        r->ctId = first.ctId;
        for (int i = 0; i < 3; ++i) {
            r->doseDims[i] = first.doseDims[i];
            r->doseRes[i] = first.doseRes[i];
            r->doseCorner[i] = first.doseCorner[i];
        }
        r->ticket = ++sv->tickets;
        r->tSubmit = now_ms();
        *ticket = r->ticket;
        sv->queue.push_back(r->ticket);
        sv->reqs[r->ticket] = std::move(r);
    }
    sv->cv.notify_all();
    return PBS_OK;
}

int pbs_server_query(pbs_server *sv, long long ticket) {
    if (!sv) return set_error(PBS_E_INVALID, "null server");
    std::lock_guard<std::mutex> g(sv->mu);
    ServerReq *r = nullptr;
    int rc = server_request_status(sv, ticket, &r);
    if (rc != PBS_OK) return rc;
    if (r->status < 0) return set_error(r->status, r->error);
    return r->status == 2 ? 1 : 0;
}

int pbs_server_wait(pbs_server *sv, long long ticket) {
    if (!sv) return set_error(PBS_E_INVALID, "null server");
    std::unique_lock<std::mutex> lk(sv->mu);
    ServerReq *r = nullptr;
    int rc = server_request_status(sv, ticket, &r);
    if (rc != PBS_OK) return rc;
    sv->doneCv.wait(lk, [&] { return r->status == 2 || r->status < 0; });
    if (r->status < 0) return set_error(r->status, r->error);
    return PBS_OK;
}

static int server_result(pbs_server *sv, long long ticket, float *dst, hipMemcpyKind kind) {
    int rc = pbs_server_wait(sv, ticket);
    if (rc != PBS_OK) return rc;
    float *dose = nullptr;
    size_t n = 0, fp = 0;
    {
        std::lock_guard<std::mutex> g(sv->mu);
        ServerReq &r = *sv->reqs[ticket];
        if (!r.dose) return set_error(PBS_E_INVALID, "result already fetched");
        dose = r.dose;
        n = r.doseN;
        fp = r.doseCharge;
        r.dose = nullptr;
        r.doseCharge = 0;
    }
    const hipError_t ce = hipMemcpyAsync(dst, dose, n * sizeof(float), kind, sv->copyStream);
    const hipError_t se = hipStreamSynchronize(sv->copyStream);
    {
        // the copy has completed: the buffer serves a later request of the same size; beyond an
        // eighth of the budget of idle buffers it is freed
        std::lock_guard<std::mutex> g(sv->mu);
        if (sv->idleBytes + fp <= sv->budget / 8) {
            sv->idleDose.emplace(fp, dose);
            sv->idleBytes += fp;
        } else {
            (void)hipFree(dose);
            sv->used -= std::min(fp, sv->used);
        }
    }
    HIP_TRY(ce);
    HIP_TRY(se);
    sv->cv.notify_all();
    return PBS_OK;
}

int pbs_server_result(pbs_server *sv, long long ticket, float *doseHost) {
    if (!sv || !doseHost) return set_error(PBS_E_INVALID, "null argument");
    return server_result(sv, ticket, doseHost, hipMemcpyDeviceToHost);
}

int pbs_server_result_device(pbs_server *sv, long long ticket, float *doseDevice) {
    if (!sv || !doseDevice) return set_error(PBS_E_INVALID, "null argument");
    return server_result(sv, ticket, doseDevice, hipMemcpyDeviceToDevice);
}

int pbs_server_timing(pbs_server *sv, long long ticket, double *queueMs, double *totalMs) {
    if (!sv) return set_error(PBS_E_INVALID, "null server");
    std::lock_guard<std::mutex> g(sv->mu);
    ServerReq *r = nullptr;
    int rc = server_request_status(sv, ticket, &r);
    if (rc != PBS_OK) return rc;
    if (queueMs) *queueMs = r->tStart > 0 ? r->tStart - r->tSubmit : -1.0;
    if (totalMs) *totalMs = r->tDone > 0 ? r->tDone - r->tSubmit : -1.0;
    return PBS_OK;
}

int pbs_server_release_state(pbs_server *sv, long long stateTicket) {
    if (!sv) return set_error(PBS_E_INVALID, "null server");
    std::lock_guard<std::mutex> g(sv->mu);
    auto it = sv->states.find(stateTicket);
    if (it == sv->states.end()) return set_error(PBS_E_INVALID, "unknown resident beam");
    for (long long q : sv->queue)
        if (sv->reqs[q]->kind == 1 && sv->reqs[q]->stateTicket == stateTicket)
            return set_error(PBS_E_INVALID, "updates of this beam are still queued");
    for (long long t : sv->slotTicket)
        if (t && sv->reqs[t]->kind == 1 && sv->reqs[t]->stateTicket == stateTicket)
            return set_error(PBS_E_INVALID, "an update of this beam is running");
    for (pbs_engine *e : sv->slots) (void)hipStreamSynchronize(e->stream);
    auto first = sv->reqs.find(stateTicket);
    const size_t b = first != sv->reqs.end() ? first->second->stateCharge : 0;
    if (first != sv->reqs.end()) first->second->stateCharge = 0;
    sv->used -= std::min(b, sv->used);
    sv->cts[sv->stateCt[stateTicket]]->users--;
    sv->states.erase(it);
    sv->stateCt.erase(stateTicket);
    return PBS_OK;
}

int pbs_server_memory(pbs_server *sv, double *usedBytes, double *budgetBytes) {
    if (!sv) return set_error(PBS_E_INVALID, "null server");
    std::lock_guard<std::mutex> g(sv->mu);
    if (usedBytes) *usedBytes = double(sv->used);
    if (budgetBytes) *budgetBytes = double(sv->budget);
    return PBS_OK;
}

}  // extern "C"
