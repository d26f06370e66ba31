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
};

// Per-beam values produced on the device.
struct BeamScalars {
    int beamFirstInside, beamFirstOutside, firstPastCutoffAll, guaranteedPassive;
    float entryZ, pxMultX, pxMultY, pad0;
    int maxPassive, kEndMax, activeLayers, pad1;
    int boxStart[3], boxEnd[3];  // transfer box [start, end) in dose index space (x, y from grid rounding)
    int boxMax[3], pad2;
};

struct LayerScalars {
    int afterLast, passiveRaw, passive, kEnd;  // kEnd: last superposed step + 1
    float entrySigX, entrySigY;
    int pad0, pad1;
};

}  // namespace pbs
