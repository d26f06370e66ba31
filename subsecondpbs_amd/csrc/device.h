// Device-side data layouts shared by the kernels and the engine (plain structs passed by
// value to kernels or stored in device memory).
#pragma once

#include <hip/hip_runtime.h>

namespace pbs {

// Fan (beam's-eye-view) geometry of one beam.
struct FanGeom {
    float g2im[9], g2imOff[3];   // gantry -> CT index (row-major)
    float d2g[9], d2gOff[3];     // dose index -> gantry
    float g2dose[9], g2doseOff[3];
    float fc[3], fd[3];          // fan corner, fan delta (mm); fd[2] < 0
    float dist[2];               // source distances (mm)
    int rx, ry;                  // traced rays
    int sx, sy;                  // rays padded to 32 x 8
    int K;                       // tracer steps
    int ctDims[3];
    int doseDims[3];
    float volConst, volLin, volSq;
    float maxCutoffWepl;
    int numLayers;
    int legacy;
};

struct LutDev {
    const float *cidd; int nE, nS;
    const float *density; int nDensity; float densityScale;
    const float *sp; int nSp; float spScale;
    const float *rrl; int nRrl; float rrlScale;
};

// Per-layer values known on the host before any kernel runs.
struct LayerHost {
    float energyIdx, scaleFact, peakDepth, rangeStop, cutoffWepl, airQuad, airLin;
    float convA, convB, sigmaX, sigmaY, pad;
    float airA, airC;  // sigmaSqAirCoefs(peakDepth): entry sigma^2 of the halo
    float pad1, pad2;
};

// Halo (nuclear) branch inputs and outputs of the fill; rayToNuc == nullptr switches it off.
struct HaloFill {
    const int *rayToNuc;       // [ry][rx]: halo cell of the spot centred on the ray, or -1
    const float *weights;      // [L][nucPlane] spot weights
    float2 *out;               // [L][K][nucPlane] (IDD, rSigmaEff)
    const float *nucWeightLut; // [nE][nS] fraction of the energy in the halo
    const float *nucSqSigmaLut;// [nE][nS] halo sigma^2
    int nucPlane;
    float spotDist;            // spot spacing in rays
    float rSpotDistSq;         // 1 / spotDist^2
};

// Per-beam values produced on the device.
struct BeamScalars {
    int beamFirstInside, beamFirstOutside, firstPastCutoffAll, guaranteedPassive;
    float entryZ, pxMultX, pxMultY, pad0;
    int maxPassive, mainEndMax, activeLayers, numGroups;
    int tailPlanes, overflow, radOverflow, pad2;  // radOverflow: a tile past the largest radius
    int boxStart[3], boxEnd[3];  // union of the group transfer boxes, [start, end)
    int boxMax[3], pad3;
};

// Per-work-group results of the one-pass trace, reduced by the upload launch that follows the
// trace (min == nullptr when the beam's trace reduced them itself).
struct TracePartials {
    const float *min;    // [K][pitch]: minimum WEPL per step and work-group (1e20: none)
    const int *fi, *fo;  // [n]: least entry step, greatest exit step per work-group
    int pitch, n;
};

struct LayerScalars {
    int afterLast, passiveRaw, passive, mainEnd;  // mainEnd: superposed into the main BEV below this step
    float entrySigX, entrySigY;
    int group, pad0;
};

// Layers with the same passive step share a transfer box (a "group"). The main BEV holds
// every layer's steps below its group's tailStart; the group's own tail BEV holds steps
// [tailStart, tailStart + tailCount), sampled only by dose voxels inside the group's box.
struct GroupScalars {
    int passive, tailStart, tailCount, tailBase;
    int boxStart[3], boxEnd[3];
    int boxMax[3], pad;
};

}  // namespace pbs
