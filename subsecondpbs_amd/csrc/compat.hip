// C++ overload with the reference's original subsecondWrapper signature
// (include/core/raytracedicom_integration.h:132-144): converts the std::vector structs to the
// POD form of pbs_abi.h and calls pbs_subsecond_wrapper.
#include <cstdio>
#include <vector>

#include "../../include/pbs_abi.h"
#include "../../include/raytracedicom_integration.h"

namespace {

template <typename T>
const T *ptr_or_null(const std::vector<T> &v) { return v.empty() ? nullptr : v.data(); }

// The array shapes a beam's vectors must agree on (wrapper:1113-1150); nullptr when they do.
const char *shape_error(const RTDBeamSettings &s) {
    const size_t L = s.energies.size();
    if (!s.layerSpotCounts.empty()) {
        if (s.layerSpotCounts.size() != L) return "layerSpotCounts needs one count per energy layer";
        long total = 0;
        for (int c : s.layerSpotCounts) total += c;
        if (!s.spotWeights.empty() && size_t(total) != s.spotWeights.size())
            return "the layer spot counts do not add up to the number of spot weights";
        if (!s.spotPositions.empty()) {
            if (s.spotPositions.size() % 2) return "spotPositions holds (x, y) pairs: its length must be even";
            if (size_t(total) != s.spotPositions.size() / 2)
                return "the layer spot counts do not add up to the number of spot positions";
        }
    }
    if (!s.spotBeamDirections.empty()) {
        if (s.spotBeamDirections.size() % 3) return "spotBeamDirections holds 3-vectors: its length must be a multiple of 3";
        if (!s.spotPositions.empty() && s.spotBeamDirections.size() / 3 != s.spotPositions.size() / 2)
            return "spotBeamDirections needs one direction per spot position";
    }
    if (!s.layerLongitudinalCutoffs.empty() && s.layerLongitudinalCutoffs.size() != L)
        return "layerLongitudinalCutoffs needs one value per energy layer";
    if (!s.profileData.empty() && s.profileSetting.size() < 3)
        return "profileData needs profileSetting = (first depth, depth step, depth count)";
    if (s.beamParaData.size() % 3) return "beamParaData holds rows of 3 values: its length must be a multiple of 3";
    return nullptr;
}

}  // namespace

extern "C" void subsecondWrapper(const float *ctData, const int3 &ctDims, const float3 &ctResolution,
                                 const float3 &ctCorner, float *doseData, const int3 &doseDims,
                                 const float3 &doseResolution, const float3 &doseCorner,
                                 const RTDBeamSettings *beamSettings, size_t numBeams,
                                 const RTDEnergyStruct *energyData, int gpuId, bool nuclearCorrection, int verbose) {
    if (!beamSettings || !energyData) {
        std::fprintf(stderr, "[RTD_ENTRY_ASSERT] beamSettings and energyData must not be null\n");
        return;
    }
    std::vector<pbs_beam_settings> beams(numBeams);
    std::vector<std::vector<float>> sigmas(numBeams);
    for (size_t i = 0; i < numBeams; ++i) {
        const RTDBeamSettings &s = beamSettings[i];
        pbs_beam_settings &b = beams[i];
        b = pbs_beam_settings{};
        b.numLayers = int(s.energies.size());
        b.energies = ptr_or_null(s.energies);
        sigmas[i].resize(2 * s.spotSigmas.size());
        for (size_t l = 0; l < s.spotSigmas.size(); ++l) {
            sigmas[i][2 * l] = s.spotSigmas[l].x;
            sigmas[i][2 * l + 1] = s.spotSigmas[l].y;
        }
        b.spotSigmas = ptr_or_null(sigmas[i]);
        b.raySpacing[0] = s.raySpacing.x; b.raySpacing[1] = s.raySpacing.y;
        b.steps = s.steps;
        b.sourceDist[0] = s.sourceDist.x; b.sourceDist[1] = s.sourceDist.y;
        b.spotOffset[0] = s.spotOffset.x; b.spotOffset[1] = s.spotOffset.y; b.spotOffset[2] = s.spotOffset.z;
        b.spotDelta[0] = s.spotDelta.x; b.spotDelta[1] = s.spotDelta.y; b.spotDelta[2] = s.spotDelta.z;
        b.maxSubspotsPerLayer = s.maxSubspotsPerLayer;
        b.subspotData = ptr_or_null(s.subspotData);
        b.layerSpotCounts = ptr_or_null(s.layerSpotCounts);
        b.numSpots = int(s.spotWeights.size());
        b.spotPositions = ptr_or_null(s.spotPositions);
        b.spotWeights = ptr_or_null(s.spotWeights);
        b.spotBeamDirections = ptr_or_null(s.spotBeamDirections);
        b.layerLongitudinalCutoffs = ptr_or_null(s.layerLongitudinalCutoffs);
        const float3 *v[4] = {&s.beamDirection, &s.bmXDirection, &s.bmYDirection, &s.sourcePosition};
        float *d[4] = {b.beamDirection, b.bmXDirection, b.bmYDirection, b.sourcePosition};
        for (int q = 0; q < 4; ++q) { d[q][0] = v[q]->x; d[q][1] = v[q]->y; d[q][2] = v[q]->z; }
        b.sad = s.sad;
        b.refPlaneZ = s.refPlaneZ;
        b.roiLinearIndices = ptr_or_null(s.roiLinearIndices);
        b.numRoiLinearIndices = int(s.roiLinearIndices.size());
        // the WEQ header in use: waterEquivalence, else its header-only alias (wrapper:999-1002)
        const std::vector<float> &weq = s.waterEquivalence.size() >= 9 ? s.waterEquivalence : s.rayWeqHeader;
        b.waterEquivalence = weq.size() >= 9 ? weq.data() : nullptr;
        b.waterEquivalenceCount = weq.size() >= 9 ? weq.size() : 0;
        b.spotPositionsAreIndices = s.spotPositionsAreIndices ? 1 : 0;
        // the profile model travels along; its hooks stay off, as in the reference (wrapper:2812-2814)
        b.profileEnergies = ptr_or_null(s.profileEnergies);
        b.numProfileEnergies = int(s.profileEnergies.size());
        b.profileData = ptr_or_null(s.profileData);
        b.profileDataCount = s.profileData.size();
        b.profileSetting = ptr_or_null(s.profileSetting);
        b.numProfileSetting = int(s.profileSetting.size());
        b.beamParaData = ptr_or_null(s.beamParaData);
        b.numBeamParaData = int(s.beamParaData.size());
        b.beamParaPos = s.beamParaPos;
        b.carbonProfileHooks = 0;
        if (s.energies.size() != s.spotSigmas.size()) {
            std::fprintf(stderr, "[RTD_ENTRY_ASSERT] beamIdx=%zu spotSigmas size must match energies\n", i);
            return;
        }
        if (s.subspotData.size() != s.energies.size() * size_t(s.maxSubspotsPerLayer) * 5) {
            std::fprintf(stderr, "[RTD_ENTRY_ASSERT] beamIdx=%zu beam.subspotData size must equal numLayers*maxSubspotsPerLayer*5\n", i);
            return;
        }
        // the vector-size checks of the entry validation (wrapper:1113-1150) that the POD form
        // cannot make, as it carries no sizes for these arrays
        const char *bad = shape_error(s);
        if (bad) {
            std::fprintf(stderr, "[RTD_ENTRY_ASSERT] beamIdx=%zu %s\n", i, bad);
            return;
        }
    }
    const RTDEnergyStruct &e = *energyData;
    if (e.ciddMatrix.size() != size_t(e.nEnergies) * size_t(e.nEnergySamples) ||
        e.energiesPerU.size() != size_t(e.nEnergies) || e.peakDepths.size() != size_t(e.nEnergies) ||
        e.scaleFacts.size() != size_t(e.nEnergies)) {
        std::fprintf(stderr, "[RTD_ENTRY_ASSERT] energy table sizes do not match nEnergies/nEnergySamples\n");
        return;
    }
    pbs_energy_struct es{};
    es.nEnergySamples = e.nEnergySamples;
    es.nEnergies = e.nEnergies;
    es.energiesPerU = ptr_or_null(e.energiesPerU);
    es.peakDepths = ptr_or_null(e.peakDepths);
    es.scaleFacts = ptr_or_null(e.scaleFacts);
    es.ciddMatrix = ptr_or_null(e.ciddMatrix);
    es.nDensitySamples = e.nDensitySamples; es.densityScaleFact = e.densityScaleFact;
    es.densityVector = ptr_or_null(e.densityVector);
    es.nSpSamples = e.nSpSamples; es.spScaleFact = e.spScaleFact; es.spVector = ptr_or_null(e.spVector);
    es.nRRlSamples = e.nRRlSamples; es.rRlScaleFact = e.rRlScaleFact; es.rRlVector = ptr_or_null(e.rRlVector);
    es.nucWeightMatrix = ptr_or_null(e.nucWeightMatrix);
    es.nucSqSigmaMatrix = ptr_or_null(e.nucSqSigmaMatrix);
    const int cd[3] = {ctDims.x, ctDims.y, ctDims.z};
    const float cr[3] = {ctResolution.x, ctResolution.y, ctResolution.z};
    const float cc[3] = {ctCorner.x, ctCorner.y, ctCorner.z};
    const int dd[3] = {doseDims.x, doseDims.y, doseDims.z};
    const float dr[3] = {doseResolution.x, doseResolution.y, doseResolution.z};
    const float dc[3] = {doseCorner.x, doseCorner.y, doseCorner.z};
    const int rc = pbs_subsecond_wrapper(ctData, cd, cr, cc, doseData, dd, dr, dc, beams.data(), numBeams, &es, gpuId,
                                         nuclearCorrection ? 1 : 0, verbose);
    if (rc != PBS_OK) std::fprintf(stderr, "[RTD_ENTRY_ASSERT] %s\n", pbs_last_error());
}

extern "C" RTDBeamSettings *createRTDBeamSettings() { return new RTDBeamSettings(); }
extern "C" void destroyRTDBeamSettings(RTDBeamSettings *beam) { delete beam; }
extern "C" RTDEnergyStruct *createRTDEnergyStruct() { return new RTDEnergyStruct(); }
extern "C" void destroyRTDEnergyStruct(RTDEnergyStruct *energy) { delete energy; }
