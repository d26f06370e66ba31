/*
 * C++ drop-in for the reference's integration header
 * (include/core/raytracedicom_integration.h:31-144): the same RTDBeamSettings and
 * RTDEnergyStruct aggregates (std::vector members, HIP float2/float3/int3 in place of the
 * CUDA vector types of identical layout) and the same subsecondWrapper signature, so code
 * written against the reference compiles and links against libpbs_hip.so unchanged.
 *
 * subsecondWrapper keeps the reference's void return: on failure it prints
 * "[RTD_ENTRY_ASSERT] <message>" to stderr and leaves doseData untouched, as the reference
 * does for validation failures (wrapper:1025-1075). pbs_subsecond_wrapper (pbs_abi.h) is
 * the status-returning C form.
 */
#ifndef RAYTRACEDICOM_INTEGRATION_H
#define RAYTRACEDICOM_INTEGRATION_H

#include <hip/hip_runtime.h>

#include <vector>

#define P_INV 0.5649718f
#define E_COEF 8.639415f
#define SQRT2 1.41421356f
#define E_REF_SQ 198.81f
#define SIGMA_DELTA 0.21f
#define RAY_WEIGHT_CUTOFF 1e-6f

struct RTDBeamSettings {
    std::vector<float> energies;
    std::vector<float2> spotSigmas;
    float2 raySpacing;
    unsigned int steps;
    float2 sourceDist;
    float3 spotOffset;
    float3 spotDelta;
    float3 gantryToImOffset;
    float3 gantryToImMatrix;
    float3 gantryToDoseOffset;
    float3 gantryToDoseMatrix;
    int maxSubspotsPerLayer = 0;
    std::vector<float> subspotData;
    std::vector<int> layerSpotCounts;
    std::vector<float> spotPositions;
    bool spotPositionsAreIndices = false;
    std::vector<float> spotWeights;
    std::vector<float> spotBeamDirections;
    std::vector<float> layerLongitudinalCutoffs;
    std::vector<float> waterEquivalence;
    std::vector<float> rayWeqHeader;
    std::vector<float> profileEnergies;
    std::vector<float> profileData;
    std::vector<float> profileSetting;
    std::vector<float> beamParaData;
    float beamParaPos = 0.0f;
    float3 beamDirection = {0.0f, 0.0f, 1.0f};
    float3 bmXDirection = {1.0f, 0.0f, 0.0f};
    float3 bmYDirection = {0.0f, 1.0f, 0.0f};
    float3 sourcePosition = {0.0f, 0.0f, 0.0f};
    float sad = 0.0f;
    float refPlaneZ = 0.0f;
    std::vector<int> roiLinearIndices;
};

struct RTDEnergyStruct {
    int nEnergySamples = 0;
    int nEnergies = 0;
    std::vector<float> energiesPerU;
    std::vector<float> peakDepths;
    std::vector<float> scaleFacts;
    std::vector<float> ciddMatrix;
    int nDensitySamples = 0;
    float densityScaleFact = 0.0f;
    std::vector<float> densityVector;
    int nSpSamples = 0;
    float spScaleFact = 0.0f;
    std::vector<float> spVector;
    int nRRlSamples = 0;
    float rRlScaleFact = 0.0f;
    std::vector<float> rRlVector;
    int nNucEnergySamples = 0;
    int nNucEnergies = 0;
    std::vector<float> nucEnergiesPerU;
    std::vector<float> nucPeakDepths;
    std::vector<float> nucScaleFacts;
    std::vector<float> nucWeightMatrix;
    std::vector<float> nucSqSigmaMatrix;
    bool nuclearTablesAlignedToPrimaryAxis = false;
    bool nuclearTablesResampledFromReference = false;
};

extern "C" void subsecondWrapper(const float *ctData, const int3 &ctDims, const float3 &ctResolution,
                                 const float3 &ctCorner, float *doseData, const int3 &doseDims,
                                 const float3 &doseResolution, const float3 &doseCorner,
                                 const RTDBeamSettings *beamSettings, size_t numBeams,
                                 const RTDEnergyStruct *energyData, int gpuId = 0, bool nuclearCorrection = false,
                                 int verbose = 0);

extern "C" RTDBeamSettings *createRTDBeamSettings();
extern "C" void destroyRTDBeamSettings(RTDBeamSettings *beam);
extern "C" RTDEnergyStruct *createRTDEnergyStruct();
extern "C" void destroyRTDEnergyStruct(RTDEnergyStruct *energy);

#endif /* RAYTRACEDICOM_INTEGRATION_H */
