// Launch interface of the gfx950 kernels (kernels.hip) used by the engine.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "device.h"

namespace pbs {

constexpr int kSideStreams = 2;  // superposition radius classes run on 1 + kSideStreams streams

// ---------------------------------------------------------------------------------------
// Multi-beam launches
// ---------------------------------------------------------------------------------------
// Every kernel of the dose chain is a body that runs one beam's grid. Called alone it is
// launched as usual. In the batched form the engine runs with a Recorder installed: each launch
// becomes a LaunchRec (the kernel's arguments, grid and LDS), and the batch executor
// (stage_records, issue_records) issues the records of one pass for all beams of the batch as a single launch in
// which every work-group finds its beam in a table of first work-groups.
struct MultiTable {
    const unsigned *base;  // [nb + 1]: first work-group of each beam's grid; base[nb] = total
    const uint4 *dims;     // [nb]: the beam's own grid (x, y, z)
    int nb;
};
using OneFn = void (*)(const void *args, dim3 grid, size_t lds, hipStream_t s);
using ManyFn = void (*)(const void *args, const MultiTable &t, unsigned blocks, size_t lds, hipStream_t s);
struct LaunchRec {
    int key;   // pass position: stage * 64 + launch within the stage; equal keys of different beams merge
    int beam;  // the beam's index in the batch
    int seq;   // the record's index within its beam
    OneFn one;
    ManyFn many;
    std::vector<unsigned char> args;  // the kernel's Args
    size_t argAlign;
    dim3 grid;
    size_t lds;
    unsigned manyBlocks;  // work-groups of the multi-beam launch; 0: the beams' grids concatenated
    const void *target;   // records that share a launch share this too (the transfers' dose grid)
    bool par;             // one of a run of launches that do not depend on each other (radius classes)
};
struct Recorder {
    int beam = 0, stage = 0, sub = 0, seq = 0;
    bool par = false;  // set around launches that do not depend on each other
    std::vector<LaunchRec> recs;
    void set_stage(int st) {
        stage = st;
        sub = 0;
    }
};
// Installed by the batch executor around the recording of one beam (nullptr: launch directly).
extern thread_local Recorder *t_rec;
// Passes that every beam's transfers share: all of them go into one launch after every other pass,
// in beam order (main grid before halo grid), so each voxel adds the beams in the order the
// one-beam-at-a-time form does.
constexpr int kTransferKey = 1 << 24;
// Bytes of the table stage_records writes for `recs` (arguments, first work-groups, grids).
size_t records_table_bytes(const std::vector<LaunchRec> &recs);
// One multi-beam launch of a batch: its kernel, where its table lies (offsets from the table's
// start: Args array, first work-groups, grids), beams, work-groups and LDS. Two batches whose
// launch lists are equal differ only in their tables' contents.
struct RecLaunch {
    ManyFn many;
    size_t off, baseOff, dimsOff;
    int nb;
    unsigned blocks;
    size_t lds;
    bool par;  // may run beside the neighbouring launches marked par
    bool operator==(const RecLaunch &o) const {
        return many == o.many && off == o.off && baseOff == o.baseOff && dimsOff == o.dimsOff && nb == o.nb &&
               blocks == o.blocks && lds == o.lds && par == o.par;
    }
};
// Side streams and events for the independent launches of a batch: a run of par launches is
// dealt round robin to the main stream and the sides (forked from and joined to the main one).
struct ForkJoin {
    hipStream_t side[kSideStreams];
    hipEvent_t fork;
    hipEvent_t join[kSideStreams];
};
// Writes the table of `recs` (at least records_table_bytes) to `hostTable` and lists the launches:
// records of equal key and kernel as one multi-beam launch, in key order. Returns the bytes.
size_t stage_records(const std::vector<LaunchRec> &recs, unsigned char *hostTable, std::vector<RecLaunch> &out);
// Issues the launches on `s`, reading the table from `devTable` (a device copy of the host table).
hipError_t issue_records(const std::vector<RecLaunch> &launches, unsigned char *devTable, hipStream_t s,
                         const ForkJoin *fj);

constexpr int kMaxRadDev = 32;
constexpr int kMinTilesDev = 16;
constexpr int kSlotPitch = 96;                    // 32 + 2 * 32 columns
constexpr int kSlotFloats = kSlotPitch * 72;      // (8 + 2 * 32) rows
constexpr int kNumClasses = 5;
constexpr int kMaxLayersDev = 256;
// transfer voxel bricks (x, y along the dose grid's cross-beam axes, z along its depth axis)
constexpr int kNumTransferShapes = 5;

struct ConvArgs {
    int rx, ry, spotX, spotY, nSub, subspotMode;
    float cpbResX, cpbResY, deltaX, deltaY, offX, offY;
    float spotOffX, spotOffY, cpbCornerX, cpbCornerY, lenToMm;
};

struct CpbArgs {
    int nx, ny, nSub, pad;
    float cx, cy, rx_, ry_;
    float bd[3], bx[3], by[3], src[3];
    float sad;
};

struct FillConsts {
    float eRefSq, sigmaDelta;  // eRefSq includes the ion's (Z/A)^2 scattering factor
    int suppressDistalDip;
    int unitWeight;  // 1: every ray weight is taken as 1 (the weight-independent field of a resident beam)
    float rangeCoef, rangeExp, restMass;  // residual energy c (peak - depth)^q (MeV/u), rest mass / nucleon
    float pad;
};

struct TransferArgs {
    float coefI[3], coefJ[3], coefOffset[3], inc[3], globalOffset[3];
    float normDist[2];
};

// Scratch floats the trace needs (the segment prefixes and carries), 8-byte aligned start.
size_t trace_scratch_floats(const FanGeom &g);
// WEQ-volume input of the trace stage: samples [nx][ny][nSteps], lattice origin / spacing, the
// ray grid's origin / spacing (same units), active steps and the step length in mm.
struct WeqArgs {
    const float *body;
    int nx, ny, nSteps, active;
    float stepMm, rayX0, rayY0, rayDx, rayDy, x0, y0, dx, dy;
};
hipError_t launch_weq_fill(const FanGeom &g, const WeqArgs &a, float *density, float *wepl, int *fi, int *fo,
                          hipStream_t s);
// The CT trace. Up to kMaxTraceSteps steps it runs in one pass whose per-beam reductions
// (weplMin per step, beamFirstInside / beamFirstOutside of bs) are left as partials in the
// scratch (*parts), for the upload launch to reduce; past that in three passes that reduce
// them themselves (parts->min == nullptr).
constexpr int kMaxTraceSteps = 512;
hipError_t launch_trace(const FanGeom &g, const float *ct, const LutDev &lut, float *density, float *wepl, int *fi,
                        int *fo, float *scratch, float *weplMin, BeamScalars *bs, TracePartials *parts,
                        hipStream_t s);
hipError_t launch_trace_reduce(const FanGeom &g, const float *wepl, const int *fi, const int *fo, float *weplMin,
                               BeamScalars *bs, hipStream_t s);
// The beam's cut-offs and entry sigmas from weplMin and the beam's entry / exit steps.
hipError_t launch_beam_setup(const FanGeom &g, const float *weplMin, const LayerHost *lh, BeamScalars *bs,
                             LayerScalars *ls, const float *subspots, int maxSub, float *subEntry, float nucFactor,
                             int *classCount, int *hist, hipStream_t s);
hipError_t launch_conv(const ConvArgs &a, const float *spot, float *interm, float *rayW, const LayerScalars *ls,
                       const BeamScalars *bs, const float *subspots, const float *subEntry, int numLayers,
                       hipStream_t s);
hipError_t launch_cpb(const CpbArgs &a, const float *sub, float *cpb, float *rays, int rnx, int rny, float rcx,
                      float rcy, float rrx, float rry, int numLayers, hipStream_t s);
hipError_t launch_fill(const FanGeom &g, const LutDev &lut, const LayerHost *lh, LayerScalars *ls,
                       const BeamScalars *bs, const float *density, const float *wepl, const float *rayW, const int *fi,
                       const int *fo, float2 *iddSigma, int *firstPassive, const FillConsts &fc, const HaloFill &hf,
                       int *tileCode, hipStream_t s);
// Weight rescale of a resident beam's unit-weight field: IDD = W x unit IDD on rays with
// W >= 1e-6 (else 0 / inf), firstPassive, the tile codes and the layer passive step, as the fill
// writes them for weights W.
hipError_t launch_rescale(const FanGeom &g, LayerScalars *ls, const BeamScalars *bs,
                          const float *rayW, const float2 *unitIdd, const int *unitFp, float2 *iddSigma,
                          int *firstPassive, int *tileCode, hipStream_t s);
// CarbonPBS profile hooks on every layer's field after the fill (wrapper:2080-2188, 4604-4832):
// bit 0 of `hooks` replaces rSigmaEff by the profile model's (multi-Gaussian sigma from the table
// at the step's WEPL, plus the beam parameters' initial width at the step's physical depth; +inf
// where IDD <= 0), bit 1 scales IDD by the table's overall-weight channel. The table is
// [rows][depthN][channels], sampled as linear texture filtering at texel centres with clamped
// addressing (bilinear over depth and row, channels exact).
struct ProfileArgs {
    const float *table;
    const float4 *layer;  // [L]: table row (fractional), beam parameters 0..2
    int rows, depthN, channels, hooks;
    float depth0, depthStep, beamParaPos, pad;
};
hipError_t launch_profile_hooks(const FanGeom &g, const ProfileArgs &a, const float *wepl, float2 *iddSigma,
                                const LayerScalars *ls, const BeamScalars *bs, hipStream_t s);
hipError_t launch_seed_grid(int numLayers, const LayerScalars *ls, const BeamScalars *bs, LayerScalars *lsOut,
                            BeamScalars *bsOut, int *hist, int *classCount, hipStream_t s);
hipError_t launch_clear_pairs(float2 *v, size_t n, hipStream_t s);

// Uploads of a beam's small host arrays in one kernel that reads the engine's mapped pinned
// staging directly (a DMA-engine copy adds ~15 us of latency to the stream per call).
constexpr int kMaxUploadSegs = 7;
struct UploadSeg {
    const uint32_t *src;  // device-visible address in mapped pinned host memory
    uint32_t *dst;
    uint32_t words;
};
struct UploadArgs {
    UploadSeg seg[kMaxUploadSegs];
    int n;
    // the one-pass trace's per-work-group partials, reduced by one more row of the launch into
    // weplMin [K] and the beam's entry / exit steps (tp.min == nullptr: nothing to reduce)
    TracePartials tp;
    float *weplMin;
    BeamScalars *bs;
    int K;
};
hipError_t launch_upload(const UploadArgs &a, hipStream_t s);
hipError_t launch_finalize_and_tiles(const FanGeom &g, const float2 *iddSigma, LayerScalars *ls, GroupScalars *gs,
                                     BeamScalars *bs, int *tailPlaneGroup, int *tailPlaneK, int tailCap, int *tileRad,
                                     int *hist, int *effRad, int *itemR, int *tailItemR, int *classCount,
                                     int2 *classList, int maxUnits, int *overflowFlag, int codesFromFill,
                                     int *planeRM, int *slotLayers, hipStream_t s);
hipError_t launch_superpose(const FanGeom &g, const float2 *iddSigma, const int *tileRad, const int *effRad,
                            const LayerScalars *ls, const GroupScalars *gs, const BeamScalars *bs,
                            int *classCount, const int2 *classList, int maxUnits, const int *slotLayers,
                            float *mainSlots, float *tailSlots, int gridCap,
                            const hipStream_t streams[1 + kSideStreams]);
hipError_t launch_merge(const FanGeom &g, const BeamScalars *bs, const int *itemR, const float *mainSlots, float *bev,
                        const int *tailItemR, const float *tailSlots, float *tailBev, int tailCap, const int *planeRM,
                        hipStream_t s);
hipError_t launch_transfer(const FanGeom &g, const TransferArgs &ta, const BeamScalars *bs, const GroupScalars *gs,
                           const float *bev, const float *tailBev, float *dose, int extentX, int extentY, int extentZ,
                           int depthAxis, int shape, hipStream_t s);

// adjoint with respect to the spot weights (adjoint.hip)
hipError_t launch_adj_zero(float *v, size_t n, hipStream_t s);
hipError_t launch_adj_transfer(const FanGeom &g, const TransferArgs &ta, const BeamScalars *bs,
                               const GroupScalars *gs, const float *gDose, size_t doseN, float *gBev,
                               float *gTailBev, size_t mainN, size_t tailN, unsigned long long *fix,
                               unsigned *maxBits, int extentX, int extentY, int extentZ, int depthAxis,
                               hipStream_t s);
hipError_t launch_adj_superpose(const FanGeom &g, const float2 *iddSigma, const float *rayW, const GroupScalars *gs,
                                const BeamScalars *bs, const int *classCount, const int2 *classList, int maxUnits,
                                const int *slotLayers, const float *gBev, const float *gTailBev, float *gI,
                                hipStream_t s);
hipError_t launch_adj_reduce(const FanGeom &g, const LayerScalars *ls, const BeamScalars *bs, const float *gI,
                             float *gW, hipStream_t s);
hipError_t launch_adj_conv(const ConvArgs &a, const float *gW, float *gInterm, float *gSpot, const LayerScalars *ls,
                           const BeamScalars *bs, const float *subspots, const float *subEntry, int numLayers,
                           hipStream_t s);

}  // namespace pbs
