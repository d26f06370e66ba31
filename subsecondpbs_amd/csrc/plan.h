// Host-side beam planning for the pencil-beam engine: units, beam basis, spot lattice,
// CPB/ray grid, fan geometry, step counts and per-layer LUT parameters.
//
// The float arithmetic follows the reference host code operation by operation
// (src/core/raytracedicom_wrapper.cu:2446-4386), because the integers it produces (grid
// sizes and corners, step counts, layer cut-offs) must come out identical.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pbs_abi.h"

namespace pbs {

constexpr int kTileX = 32;       // SUPERP_TILE_X, Macro.cuh:58
constexpr int kTileY = 8;        // SUPERP_TILE_Y, Macro.cuh:59
constexpr int kMaxRad = 32;      // MAX_SUPERP_RADIUS, Macro.cuh:57
constexpr int kMinTilesInBatch = 16;  // MIN_TILES_IN_BATCH, Macro.cuh:60
constexpr int kMaxLayers = 256;

struct F3 { float x, y, z; };
struct M3 { F3 r0, r1, r2; };
struct Aff { M3 m; F3 v; };

F3 f3(float x, float y, float z);
F3 apply(const Aff &t, F3 a);
Aff inverse(const Aff &t);
Aff concat(const Aff &first, const Aff &second);

// Per-layer parameters consumed by device kernels (all plain floats/ints).
struct LayerParams {
    float energyIdx;       // fractional row of the CIDD table
    float scaleFact;       // CIDD samples per mm
    float peakDepth;       // mm
    float rangeStop;       // 1.05 * peak
    float cutoffWepl;      // threshold for findFirstLarger(weplMin, .)
    float airQuad, airLin; // sigma^2 air-divergence recurrence coefficients
    float convA, convB;    // sigmaSqAirCoefs at the conversion peak (conv stage)
    float sigmaX, sigmaY;  // spot sigma at iso (mm)
    float airA, airC;      // sigmaSqAirCoefs(peakDepth) (halo entry sigma)
    float pad0, pad1;
};

struct BeamPlan {
    std::string error;
    bool legacyPath = false;
    float lenToMm = 1.f, energyDepthToMm = 1.f;
    F3 bmX{}, bmY{}, bmZ{}, iso{};
    Aff g2w{}, g2im{}, g2dose{}, dose2g{};
    float sourceDist[2] = {0, 0};
    float sadInput = 0.f;           // SAD in input units (legacy CPB geometry)
    // spot lattice (lattice path)
    int spotDims[2] = {0, 0};
    float spotOffset[2] = {0, 0}, spotDelta[2] = {0, 0};
    std::vector<float> denseWeights;  // [L][sy][sx]
    std::vector<int> spotCell;        // per spot: its cell (l * sy + iy) * sx + ix in denseWeights, or -1
    bool placementPending = false;    // the lattice's axes are set, its spots not placed yet
    // grids
    float cpbCorner[3] = {0, 0, 0}, cpbRes[3] = {0, 0, 0};
    int rayDims[2] = {0, 0};
    int superpDims[2] = {0, 0};
    int steps = 0;
    float fanCorner[3] = {0, 0, 0}, fanDelta[3] = {0, 0, 0};
    float volConst = 0, volLin = 0, volSq = 0;
    float maxCutoffWepl = 0;  // 1.05 * peak depth of the highest-energy layer
    int numLayers = 0;
    std::vector<LayerParams> layers;
    bool nuclear = false;
    bool subspotConv = false;
    // WEQ-volume input (wrapper:1923-1988, 3305-3370, 3566-3640): the tracer is replaced by the
    // beam's water-equivalent samples [x][y][k]; lattice of the samples in ray-grid units
    bool weqVolume = false;
    bool spotIndices = false;  // spot positions are WEQ-lattice texel coordinates
    int weqDims[3] = {0, 0, 0};  // x_n, y_n, depth_n
    float weqStepMm = 0.f, weqX0 = 0.f, weqY0 = 0.f, weqDx = 0.f, weqDy = 0.f;
    // CarbonPBS profile hooks that apply to this beam (bit 0 sigma override, bit 1 overall
    // weight; wrapper:2816-2862, 4577-4832): the table's shape, its depth axis and, per layer,
    // the table row (fractional) and the beam parameters interpolated at it
    int profileHooks = 0;
    int profileRows = 0, profileDepthN = 0, profileChannels = 0;
    float profileDepth0 = 0.f, profileDepthStep = 0.f, beamParaPos = 0.f;
    std::vector<float> profileLayer;  // [L][4]: row index, beam parameters 0..2
};

// Explicit CPB grid in the gantry frame (used by cudaFinalDose, whose CPB grid comes from
// the ROI box rather than from the dose-box projection).
struct CpbOverride {
    float corner[2], res[2];
    int dims[2];
};

// Halo (nuclear) lattice plan of a beam (wrapper:804-929; spot spacing wrapper:4062-4100,
// 4149-4172, 4338-4350): the halo grid is the spot lattice rounded up to 32x8; the ray at a
// spot's gantry position carries that spot's halo; spotDist = spot spacing / fan ray spacing.
struct HaloPlan {
    int nucDims[2] = {0, 0};
    int rawDims[2] = {0, 0};       // the halo lattice (spot cells) before 32x8 padding
    float offset[2] = {0.f, 0.f};  // gantry position of cell (0, 0), mm
    float delta[2] = {0.f, 0.f};   // cell spacing, mm
    bool latticeFallback = false;  // index spots: the WEQ header lattice stood in for the PB grid
    float spotDist = 0.f;
    std::vector<int> rayToNuc;   // [ry][rx], -1 = no spot centred on the ray
    std::vector<float> weights;  // [L][nucY][nucX]
};
// Returns false with `error` set when the beam cannot drive the halo branch.
bool plan_halo(const BeamPlan &plan, const pbs_beam_settings &beam, HaloPlan &halo, std::string &error);

// The reference's environment overrides of its unit and input-type heuristics, read once by
// read_plan_env() when an entry point starts (an engine reads them when it is created);
// nothing here writes the environment.
//  * RTD_CT_INPUT_TYPE = hu | hu+1000 (huplus1000, hup1000) | density (rho) | spr
//    (stoppingpower, stopping_power), any case; unset or other = auto (wrapper:2288-2299);
//  * RTD_ENERGY_DEPTH_SCALE = <float>: the LUT depth unit in mm (0 or negative: the CT's
//    length unit); else RTD_ENERGY_DEPTH_UNIT = cm | mm (other values: the CT's length unit);
//    with neither, the cm-CT / mm-table heuristic decides (wrapper:2742-2790);
//  * RTD_CPB_COVERAGE_CM = <float > 0>: the legacy CPB grid's coverage in input units on both
//    axes (wrapper:3193-3202).
struct PlanEnv {
    int ctInputType = -1;          // -1 auto, 0 HU, 1 HU+1000, 2 density, 3 SPR
    int energyDepthMode = 0;       // 0 heuristic, 1 fixed energyDepthToMm, 2 the CT's length unit
    float energyDepthToMm = 0.0f;
    float cpbCoverage = 0.0f;      // > 0: legacy CPB coverage override
};
PlanEnv read_plan_env();

// Returns false (and sets plan.error) on a contract failure. With deferPlacement, a beam of
// physical spot positions is planned from its lattice's axes alone (the per-spot placement,
// most of the planning time, is left for place_plan_spots; plan.placementPending): the grids
// are the lattice path's, valid if every weighted spot then lands on a lattice point.
bool plan_beam(const int ctDims[3], const float ctRes[3], const float ctCorner[3],
               const int doseDims[3], const float doseRes[3], const float doseCorner[3],
               const pbs_beam_settings &beam, const pbs_energy_struct &energy, BeamPlan &plan,
               const CpbOverride *cpbOverride = nullptr, const PlanEnv *env = nullptr,
               bool deferPlacement = false);
// The deferred placement of a plan made with deferPlacement: the dense lattice weights and the
// spots' cells. False when a weighted spot is off the lattice: the beam is then not a lattice
// beam, and must be planned again without deferral (which takes the legacy CPB path).
bool place_plan_spots(const pbs_beam_settings &beam, BeamPlan &plan);

// findDecimalOrdered (include/utils/vector_find.h:130-145) and the wrapper's own energy
// bracket (wrapper:3926-3992).
float find_decimal_ordered(const float *v, int n, float value);
float layer_energy_index(const pbs_energy_struct &e, float energy);

// Reference CT auto-detection on a host buffer: 0 = HU, 1 = HU+1000, 2 = density/SPR.
int detect_ct_type(const float *ct, size_t n);
// convertCtToHUPlus1000 (wrapper:2337-2441) for type 0 (HU), 2 (density) or 3 (SPR): HU+1000 into
// out[n]. Density and SPR are inverted through `lut` (the density or SP table, index =
// HU+1000 x scale) by the reference's bisection, which picks one crossing where the table is not
// monotonic. Returns false when the table is empty or its scale is 0.
bool ct_to_hu_plus_1000(const float *in, size_t n, int type, const float *lut, int lutN, float scale, float *out);

// Radius bins -> launch radii (complete_superposition.cu:413-490).
void batch_bins(const int *hist34, int *batched33, int *effRad34);

}  // namespace pbs
