/*
 * C ABI of the MI355X pencil-beam final-dose engine (libpbs_hip.so).
 *
 * Plain pointers and sizes only; no HIP, torch or STL types cross this boundary. Every
 * entry point returns 0 on success or a negative PBS_E_* code, and leaves a message for
 * pbs_last_error() (thread-local). Nothing here calls exit().
 *
 * Which reference interface each entry point replaces is given next to it. All lengths
 * are millimetres unless the reference's own unit heuristics say otherwise (a CT
 * resolution below 0.3 marks centimetre input, wrapper:2728-2729).
 */
#ifndef PBS_ABI_H
#define PBS_ABI_H

#include <stddef.h>
#include "pbs_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PBS_OK 0
#define PBS_E_INVALID -1   /* argument validation failed (RTD_ENTRY_ASSERT class) */
#define PBS_E_GEOMETRY -2  /* grid / lattice / halo contract failure (reference throws) */
#define PBS_E_DEVICE -3    /* HIP runtime error */
#define PBS_E_NOMEM -4     /* device allocation failed or admission refused */
#define PBS_E_UNSUPPORTED -5

/*
 * Beam description: a POD mirror of RTDBeamSettings (include/core/raytracedicom_integration.h:31-95)
 * with each std::vector replaced by a pointer + count. Optional arrays may be NULL.
 */
typedef struct pbs_beam_settings {
    int numLayers;
    const float *energies;        /* [numLayers] MeV/u */
    const float *spotSigmas;      /* [numLayers][2] sigma_x, sigma_y at iso in air */
    float raySpacing[2];
    unsigned int steps;           /* kept for parity; the tracer derives its own step count */
    float sourceDist[2];
    float spotOffset[3];
    float spotDelta[3];
    int maxSubspotsPerLayer;
    const float *subspotData;     /* [numLayers][maxSubspotsPerLayer][5]: dx, dy, w, sx, sy */
    const int *layerSpotCounts;   /* [numLayers] or NULL (then the subspot->CPB path is used) */
    int numSpots;
    const float *spotPositions;   /* [numSpots][2] gantry-frame x, y on the iso plane */
    const float *spotWeights;     /* [numSpots] */
    const float *spotBeamDirections; /* [numSpots][3] or NULL */
    const float *layerLongitudinalCutoffs; /* [numLayers] or NULL (audit only, as in the reference) */
    float beamDirection[3];
    float bmXDirection[3];
    float bmYDirection[3];
    float sourcePosition[3];
    float sad;
    float refPlaneZ;
    const int *roiLinearIndices;  /* x-major x*ny*nz + y*nz + z, or NULL */
    int numRoiLinearIndices;
    int useSubspotConvolution;    /* 1: per-subspot lattice convolution (reference branch
                                     wrapper:3770-3822, compiled in but switched off there) */
    /* Water-equivalent input of the CarbonPBS contract (RTDBeamSettings::waterEquivalence and
     * rayWeqHeader, raytracedicom_integration.h:63-70): a 9-value header [depth_start,
     * depth_step, depth_n, y_start, y_step, y_n, x_start, x_step, x_n] optionally followed by
     * x_n * y_n * depth_n water-equivalent depths at index (x * y_n + y) * depth_n + k (values past
     * that count are ignored, wrapper:3566-3590). With samples, they replace the CT tracer
     * (fillBevFromWeqVolumeKernel, wrapper:1923-1988), the step length and count come from the
     * header and the start plane is aligned to depth_start (wrapper:3305-3370). NULL / 0: CT
     * tracer. These members sit at the end so that zero-initialised older callers keep working. */
    const float *waterEquivalence;
    size_t waterEquivalenceCount;
    int spotPositionsAreIndices;  /* 1: spotPositions are WEQ-lattice texel coordinates (+0.5 at
                                     texel centres), rasterised bilinearly onto that lattice
                                     (wrapper:931-968, 1295-1364); needs the header */
    /* CarbonPBS lateral-profile model (RTDBeamSettings::profileEnergies, profileData,
     * profileSetting, beamParaData, beamParaPos; raytracedicom_integration.h:53-58) and its two
     * hooks on each layer's IDD / sigma field (wrapper:2080-2188, 4604-4832). profileData is
     * [rows][depth][channels] (channels = nGauss weights, nGauss sigmas, 1 overall weight),
     * profileSetting = [depth0, depthStep, depthN], beamParaData = [rows][3]. The reference
     * compiles both hooks but switches them off (wrapper:2812-2814); carbonProfileHooks turns
     * them on: bit 0 = sigma from the profile model, bit 1 = the overall-weight channel on the
     * IDD. 0 (the default) keeps the reference's behaviour and ignores the arrays. */
    const float *profileEnergies;
    int numProfileEnergies;
    const float *profileData;
    size_t profileDataCount;
    const float *profileSetting;
    int numProfileSetting;
    const float *beamParaData;
    int numBeamParaData;
    float beamParaPos;
    int carbonProfileHooks;
} pbs_beam_settings;

/* Energy/LUT payload: POD mirror of RTDEnergyStruct (raytracedicom_integration.h:97-128). */
typedef struct pbs_energy_struct {
    int nEnergySamples;
    int nEnergies;
    const float *energiesPerU;    /* [nEnergies] */
    const float *peakDepths;      /* [nEnergies] */
    const float *scaleFacts;      /* [nEnergies] */
    const float *ciddMatrix;      /* [nEnergies][nEnergySamples] */
    int nDensitySamples; float densityScaleFact; const float *densityVector;
    int nSpSamples; float spScaleFact; const float *spVector;
    int nRRlSamples; float rRlScaleFact; const float *rRlVector;
    const float *nucWeightMatrix;   /* [nEnergies][nEnergySamples] or NULL */
    const float *nucSqSigmaMatrix;  /* [nEnergies][nEnergySamples] or NULL */
    /* Ion of the table set (not in RTDEnergyStruct, which is proton only): the fill's residual
     * energy E = ionRangeCoef (peak - depth)^ionRangeExp in MeV/u, the rest mass per nucleon in
     * MeV, and the factor (Z/A)^2 on the multiple-scattering constant. 0 = the proton values
     * of idd_sigma.cu:172-181 (8.639415, 0.5649718, 938.3, 1). */
    float ionRangeCoef, ionRangeExp, ionRestMass, ionScatterScale;
} pbs_energy_struct;

/*
 * Whole-plan dose on the host, accumulated into doseData (the reference uploads doseData,
 * adds every beam's dose and copies it back: wrapper:2653-2662, 5642).
 * Replaces extern "C" subsecondWrapper (include/core/raytracedicom_integration.h:132-138,
 * src/core/raytracedicom_wrapper.cu:2446).
 */
int pbs_subsecond_wrapper(const float *ctData, const int ctDims[3], const float ctResolution[3],
                          const float ctCorner[3], float *doseData, const int doseDims[3],
                          const float doseResolution[3], const float doseCorner[3],
                          const pbs_beam_settings *beams, size_t numBeams,
                          const pbs_energy_struct *energy, int gpuId, int nuclearCorrection, int verbose);

/* Message of the last failing call on this thread (empty string if none). */
const char *pbs_last_error(void);
/* sizeof of the ABI structs as compiled into the library (bindings check their mirrors against
 * it): 0 pbs_beam_settings, 1 pbs_energy_struct, 2 pbs_plan_info; -1 for another value. No GPU
 * needed. */
long pbs_abi_sizeof(int which);

/* Number of visible HIP devices (0 when none; never fails). */
int pbs_device_count(void);

/* ---------------------------------------------------------------------------------------
 * Resident engine: CT, LUTs and work buffers stay in HBM between calls (batching, serving,
 * differentiation). No reference counterpart; these are the building blocks of the
 * reference's per-call flow split into reusable parts.
 * ------------------------------------------------------------------------------------- */
typedef struct pbs_engine pbs_engine;

pbs_engine *pbs_engine_create(int gpuId, const pbs_energy_struct *energy);
void pbs_engine_destroy(pbs_engine *e);
/* Upload a host CT (ctIsDevice = 0) or copy a device CT (ctIsDevice = 1); values are
 * converted to HU+1000 with the reference's auto-detection (wrapper:2290-2441).
 * ctIsDevice = 2 borrows a device buffer that already holds HU+1000 (no copy; it must outlive
 * the beams queued on it), which lets several engines share one resident CT. */
int pbs_engine_set_ct(pbs_engine *e, const float *ct, const int ctDims[3], const float ctRes[3],
                      const float ctCorner[3], int ctIsDevice);
/* Dose grid geometry used by subsequent beam calls. */
int pbs_engine_set_dose_grid(pbs_engine *e, const int doseDims[3], const float doseRes[3],
                             const float doseCorner[3]);
/* Accumulate one beam's dose into doseDevice ([nz][ny][nx] float32 device buffer) on the
 * engine's stream; returns after the work is queued (synchronize with pbs_engine_sync). */
int pbs_engine_beam_dose(pbs_engine *e, const pbs_beam_settings *beam, float *doseDevice);
/* pbs_engine_beam_dose without any device allocation. No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Every per-beam work buffer is carved from a static 256 MiB device scratch
 * (one per device, held from the library's load until the process ends) and the beam's small
 * arrays go up through it when the engine's pinned staging is too small; no hipMalloc, hipFree or
 * host-pinned allocation is made. Buffers the engine already holds are used when they are large
 * enough. A beam whose work buffers exceed the scratch fails with PBS_E_NOMEM and a message (the
 * config-1-size beams fit: 64^3 CT, 96x80 rays, 64 steps, 5 layers; a head-size beam does not).
 * Every call on one device shares that device's scratch, which keeps the last call's contents:
 * calls that overlap on one GPU (from several threads or engines) race, so callers serialise
 * them. Results equal pbs_engine_beam_dose's bitwise. Stage profiling must be off. */
int pbs_engine_beam_dose_noalloc(pbs_engine *e, const pbs_beam_settings *beam, float *doseDevice);
/* Adjoint with respect to the spot weights. No counterpart in the reference, which has no
 * gradient path; it serves weight optimisation (the transposed superposition gather, then the
 * transposed subspot convolution). For J(dose) with dJ/dDose = gradDoseDevice ([nz][ny][nx]
 * float32 device buffer on this engine's dose grid) it writes gradSpotWeights[numSpots] (host)
 * = dJ/dw at the beam's weights, holding live masks, tile radii and step ranges fixed (the dose
 * is linear in the weights for fixed masks). Spot-lattice beams only (spot positions and
 * weights given), halo off; returns after the gradient is on the host. The BEV gradient is
 * accumulated with float atomics, so results agree run to run to rounding, not bitwise. */
int pbs_engine_beam_adjoint(pbs_engine *e, const pbs_beam_settings *beam, const float *gradDoseDevice,
                            float *gradSpotWeights);
/* Resident beams on a plain engine: no counterpart in the reference; they serve callers that
 * re-evaluate one beam at many spot weights (weight optimisation, the served form's updates).
 * pbs_engine_keep_beam runs the beam into doseDevice (accumulated) and keeps its weight-
 * independent state in HBM: plan, fan geometry, post-trace scalars and the unit-weight IDD /
 * rSigmaEff field. pbs_engine_update_beam adds the beam's dose at new weights (numSpots, the
 * beam's spot order) to doseDevice, re-running only the convolution, the rescale of the unit
 * field, the superposition, the merge and the transfer. pbs_engine_update_adjoint gives
 * dJ/dw at the new weights from the same state (no trace, no fill). A resident beam is used
 * with the engine that made it, whose CT, dose grid and nuclear model (off) must stay as they
 * were. Spot-lattice beams only, halo off. Results use IDD = W x unit IDD, one rounding from the
 * plain form's dose.
 * pbs_resident_beam_use_graph(rb, 1): updates are replayed as one captured HIP graph while the
 * engine, dose buffer, dose grid, launch shapes and device buffers are those captured (the
 * first call, and any call after a change, runs eagerly and captures again). */
typedef struct pbs_resident_beam pbs_resident_beam;
int pbs_engine_keep_beam(pbs_engine *e, const pbs_beam_settings *beam, float *doseDevice, pbs_resident_beam **out);
int pbs_engine_update_beam(pbs_engine *e, pbs_resident_beam *rb, const float *spotWeights, int numSpots,
                           float *doseDevice);
int pbs_engine_update_adjoint(pbs_engine *e, pbs_resident_beam *rb, const float *spotWeights, int numSpots,
                              const float *gradDoseDevice, float *gradSpotWeights);
int pbs_resident_beam_use_graph(pbs_resident_beam *rb, int enable);
void pbs_resident_beam_destroy(pbs_resident_beam *rb);
int pbs_engine_sync(pbs_engine *e);
/* Raw stream handle (hipStream_t) the engine queues work on. */
void *pbs_engine_stream(pbs_engine *e);

/* Per-stage device timing with HIP events on the engine stream (off by default). */
int pbs_engine_set_profiling(pbs_engine *e, int enable);
/* Device time per stage summed over the beams queued since the last call (profiling on):
 * msOut[8] = trace, conv, fill, tiles, superpose, merge, transfer, halo. */
int pbs_engine_stage_times(pbs_engine *e, float *msOut, int *beams);
/* Launch shapes of this engine, chosen by the autotuner (subsecondpbs_amd/autotune.py); no
 * counterpart in the reference. None of them changes a result: every unit of work is computed
 * the same way whatever block runs it. Keys:
 *   PBS_TUNE_SUPERPOSE_GRID  work-groups per superposition radius class (64..65536, default 4096)
 *   PBS_TUNE_TRANSFER_SHAPE  transfer voxel brick, 0..4: (8,8,4) (16,16,1) (16,8,2) (32,8,1)
 *                            (8,4,8) along (cross, cross, depth) axes of the dose grid (default 0)
 *   PBS_TUNE_ONE_STREAM      1: radius classes on the engine stream only; 0: on side streams too
 *                            (engines of pbs_batch / pbs_server have no side streams) */
#define PBS_TUNE_SUPERPOSE_GRID 1
#define PBS_TUNE_TRANSFER_SHAPE 2
#define PBS_TUNE_ONE_STREAM 3
int pbs_engine_set_tuning(pbs_engine *e, int key, int value);
int pbs_engine_get_tuning(pbs_engine *e, int key, int *value);

/* Nuclear (halo) branch for the following beams: 0 off, 1 GAUSS_FIT, 2 SOUKUP, 3 FLUKA
 * (the reference's NUCLEAR_CORR build modes, idd_sigma.cu:108-121). Requires the energy
 * struct's nuclear tables. The halo of each spot is computed on the spot lattice
 * (raytracedicom_wrapper.cu:804-929, 4062-4470), superposed on its own grid with radius
 * overflow as an error (complete_superposition.cu:558-708, reported by pbs_engine_sync) and
 * transferred with its own fan geometry (prim_transf_kernel.cu:46-78, wrapper:5510-5600). */
int pbs_engine_set_nuclear(pbs_engine *e, int model);
/* Halo grid of the last beam: nucDims[2], spotDist, (IDD, rSigmaEff) [L][K][nucY][nucX][2],
 * radius histograms [L][34], sticky overflow flag. Any pointer may be NULL. */
int pbs_engine_debug_halo(pbs_engine *e, int nucDims[2], float *spotDist, float *haloIddSigma,
                          int *haloRadHist, int *radOverflow);
/* The last beam's batched launch radius of every radius bin, [numLayers][34] (the batching of
 * complete_superposition.cu:482-490 applied on the device: bin r's tiles run at radius
 * effRad[r]); haloEffRad likewise for the halo grid (NULL to skip; the last beam must have
 * used the halo branch). For the parity tests. */
int pbs_engine_debug_bins(pbs_engine *e, int *effRad, int *haloEffRad);

/* Host planning of one beam, without a device (no reference counterpart: it exposes the
 * planner that pbs_engine_beam_dose runs first, so that tests on a machine without a GPU can
 * check it against the FP64 oracle's plan). The grids and depths follow the wrapper's host code
 * (raytracedicom_wrapper.cu:2728-3370, 3746-4329). */
typedef struct pbs_plan_info {
    int legacyPath, spotIndices, weqVolume, numLayers;
    int rayDims[2], superpDims[2], spotDims[2], steps;
    float spotOffset[2], spotDelta[2], cpbCorner[2], cpbRes[2];
    float fanCorner[3], fanDelta[3], sourceDist[2];
    float volConst, volLin, volSq;
    float lenToMm, energyDepthToMm; /* input length unit and LUT depth unit, in mm */
    int profileHooks;               /* hooks that apply (bit 0 sigma, bit 1 weight) */
    int profileRows, profileDepthN, profileChannels;
} pbs_plan_info;
/* layerEnergyIdx / layerPeakDepth [numLayers], denseWeights [numLayers][spotDims1][spotDims0]
 * (denseCap floats available), spotCells [numSpots] (the dense cell of each spot, -1 off
 * the lattice) and layerProfile [numLayers][4] (profile row index and interpolated beam
 * parameters of each layer) may be NULL. */
int pbs_plan_debug(const int ctDims[3], const float ctRes[3], const float ctCorner[3], const int doseDims[3],
                   const float doseRes[3], const float doseCorner[3], const pbs_beam_settings *beam,
                   const pbs_energy_struct *energy, pbs_plan_info *info, float *layerEnergyIdx,
                   float *layerPeakDepth, float *denseWeights, size_t denseCap, int *spotCells,
                   float *layerProfile);
/* The engine's CT adapter on a host buffer, without a device (no reference counterpart: it exposes
 * what pbs_engine_set_ct runs first, for tests on a machine without a GPU). The input type is
 * RTD_CT_INPUT_TYPE's, else detected from sampled values (raytracedicom_wrapper.cu:2288-2335);
 * HU, density and SPR are converted to HU+1000 (wrapper:2337-2441), HU+1000 is copied. out[n];
 * *typeUsed (may be NULL) = 0 HU, 1 HU+1000, 2 density, 3 SPR. */
int pbs_ct_to_hu_plus_1000(const float *ct, size_t n, const pbs_energy_struct *energy, float *out, int *typeUsed);

/* Intermediate arrays of the last pbs_engine_beam_dose call copied to host (parity tests).
 * Any pointer may be NULL. Shapes: density/wepl [K][ry][rx]; firstInside/firstOutside
 * [ry][rx]; weplMin [K]; rayWeights [L][ry][rx]; iddSigma [L][K][sy][sx][2] (IDD, rSigmaEff;
 * only steps [beamFirstInside, afterLast_l) are defined); firstPassive [L][ry][rx];
 * tileRad [L][K][sy/8][sx/32] (-1 where not computed); radHist [L][34]; bev
 * [bevZ][sy+64][sx+64] (main BEV; only planes [beamFirstInside, mainEndMax) are defined, the
 * transfer reads the others as 0); ints: see pbs_debug_ints. */
typedef struct pbs_debug_ints {
    int rayDims[2], superpDims[2], steps, numLayers, legacyPath;
    int beamFirstInside, beamFirstOutside, firstPastCutoffAll, guaranteedPassive;
    int afterLast[256], layerPassive[256];
    int transferStart[3], transferMax[3];
    float cpbCorner[3], fanCorner[3], fanDelta[3], sourceDist[2];
    float energyIdx[256], scaleFact[256], peakDepth[256];
    /* layer groups (layers sharing a passive step share a transfer box): superposed into
     * the main BEV below mainEnd; group g's tail planes [tailStart, tailStart + tailCount) */
    int numGroups, tailPlanes, tailOverflow, mainEndMax;
    int mainEnd[256], layerGroup[256];
    int groupPassive[256], groupTailStart[256], groupTailCount[256];
} pbs_debug_ints;
int pbs_engine_debug_fetch(pbs_engine *e, pbs_debug_ints *ints, float *density, float *wepl,
                           int *firstInside, int *firstOutside, float *weplMin, float *rayWeights,
                           float *iddSigma, int *firstPassive, int *tileRad, int *radHist, float *bev);

/* ---------------------------------------------------------------------------------------
 * Batched and served forms (one GPU). A batch owns `lanes` engines (own stream and
 * workspaces) that share one resident CT and LUT set; beams are dealt to lanes in submission
 * order and their kernels overlap on the device. No counterpart in the reference, which
 * computes one beam at a time on the default stream.
 * ------------------------------------------------------------------------------------- */
typedef struct pbs_batch pbs_batch;
pbs_batch *pbs_batch_create(int gpuId, const pbs_energy_struct *energy, int lanes);
void pbs_batch_destroy(pbs_batch *b);
int pbs_batch_set_ct(pbs_batch *b, const float *ct, const int ctDims[3], const float ctRes[3],
                     const float ctCorner[3], int ctIsDevice);
int pbs_batch_set_dose_grid(pbs_batch *b, const int doseDims[3], const float doseRes[3],
                            const float doseCorner[3]);
int pbs_batch_set_nuclear(pbs_batch *b, int model);
/* Sum of the doses of `numBeams` beams, accumulated into doseDevice. Spot-lattice beams on the
 * CT run as one multi-beam pass per stage of the chain (trace, upload, setup, convolution, fill,
 * tiles, each radius class, merge; the transfers last, adding the beams voxel by voxel in beam
 * order), so the dose equals the one-beam-at-a-time sum bitwise. A batch holding any other beam
 * (adjoint, resident, profile hooks, WEQ volume, legacy CPB) falls back to the lanes: lane i
 * takes beams i, i + lanes, ... into its own buffer and the lanes are added in lane order.
 * Returns when the work is queued; pbs_batch_sync waits. */
int pbs_batch_dose(pbs_batch *b, const pbs_beam_settings *beams, size_t numBeams, float *doseDevice);
/* Another energy-table set (e.g. carbon beside protons) with its nuclear model (0 off, 1
 * GAUSS_FIT, 2 SOUKUP, 3 FLUKA); returns its id (>= 1; the creation's set is 0) or a negative
 * status. */
int pbs_batch_add_table(pbs_batch *b, const pbs_energy_struct *energy, int nuclearModel);
/* pbs_batch_dose's multi-beam form with each beam's table set (tableIds NULL: all 0); beams of
 * every table set share the passes. PBS_E_UNSUPPORTED for a beam the passes do not cover. */
int pbs_batch_dose_tables(pbs_batch *b, const pbs_beam_settings *beams, const int *tableIds, size_t numBeams,
                          float *doseDevice);
/* Launches the last multi-beam call issued (0 when the lanes ran it). The launches of a
 * multi-beam call form one HIP graph, captured when the call's launch list is new and replayed
 * while later calls repeat it; pbs_batch_graph_captures counts the captures. */
int pbs_batch_last_launches(pbs_batch *b);
int pbs_batch_graph_captures(pbs_batch *b);
/* Serving: queue one beam whose dose is accumulated into doseDevice, on the next lane (round
 * robin); a doseDevice that a queued request still writes goes to that request's lane, so the
 * two accumulate in order. *ticket identifies the request for pbs_batch_query / wait. */
int pbs_batch_submit(pbs_batch *b, const pbs_beam_settings *beam, float *doseDevice, long long *ticket);
/* 1 when the request's work has finished on the device, 0 while it runs. */
int pbs_batch_query(pbs_batch *b, long long ticket);
int pbs_batch_wait(pbs_batch *b, long long ticket);
int pbs_batch_sync(pbs_batch *b);
int pbs_batch_lanes(pbs_batch *b);

/* ---------------------------------------------------------------------------------------
 * cudaFinalDose with the reference's name and argument order
 * (include/algorithms/cuda_final_dose.h:42-56, src/algorithms/cuda_final_dose.cu:230-390).
 * subspotData is a float buffer [numLayers][maxSubspotsPerLayer][5] (host or device,
 * detected) instead of a texture object; roiIndices and finalDose are device pointers as in
 * the reference. The reference stops after the CPB convolution; this one completes the
 * chain using the CT, LUTs and per-layer energies bound with pbs_final_dose_bind().
 * Returns 0 or a PBS_E_* code (PBS_OK with an untouched dose when the beam misses the ROI).
 * ------------------------------------------------------------------------------------- */
int pbs_final_dose_bind(pbs_engine *e, const float *layerEnergies, const float *layerSigmas,
                        int numLayers, float raySpacing);
int cudaFinalDose(const float *subspotData, const vec3i *roiIndices, int numROIIndices,
                  vec3f doseGridResolution, vec3f doseGridCorner, vec3f beamDirection,
                  vec3f sourcePosition, float sad, float refPlaneZ, int numLayers,
                  int maxSubspotsPerLayer, float *finalDose, vec3i doseGridDims);
/* ROI range / intersection / CPB grid of cudaFinalDose on the host (bit-exact integers;
 * cuda_final_dose.cu:66-129, 132-152, 302-308). Returns 1 if the beam meets the ROI. */
int pbs_final_dose_roi(const vec3i *roiIndicesHost, int numROIIndices, vec3f doseGridResolution,
                       vec3f doseGridCorner, vec3f beamDirection, vec3f sourcePosition, int numLayers,
                       vec3f *roiMinCorner, vec3f *roiMaxCorner, vec3i *roiDims, CPBGrid *cpbGrid);

/* ---- served form -------------------------------------------------------------------
 * No counterpart in the reference (its entry is synchronous, single request, not reentrant).
 * A request server for dose services: resident CTs, a pool of engine slots (each with its own
 * stream), and a worker thread that admits queued requests onto free slots while the device
 * memory they need (dose buffer, resident beam state, growth of the slot's work buffers) fits
 * the budget, so requests that arrive while others run overlap on the device (continuous
 * batching). A full request may keep its beam resident (keepState): its unit-weight IDD /
 * rSigmaEff field, geometry and per-layer values stay in HBM, and weight-only updates of that
 * beam re-run only the convolution, the rescale, the superposition, the merge and the transfer.
 * Every request computes on its own engine and buffers: its result does not depend on what
 * else is in flight. Results of resident beams use IDD = W x unit IDD, which can differ from the
 * plain form's single product by one rounding (float ulps of the dose).
 */
typedef struct pbs_server pbs_server;
/* slots 1..16; memoryBudgetBytes <= 0: 80 % of the free device memory at creation. */
pbs_server *pbs_server_create(int gpuId, const pbs_energy_struct *energy, int slots, double memoryBudgetBytes);
void pbs_server_destroy(pbs_server *s);
/* A resident CT (HU or HU+1000, host or device pointer, copied); *ctId names it in requests. */
int pbs_server_add_ct(pbs_server *s, const float *ct, const int ctDims[3], const float ctRes[3],
                      const float ctCorner[3], int ctIsDevice, int *ctId);
int pbs_server_remove_ct(pbs_server *s, int ctId);
/* Queue a full beam on a resident CT and dose grid; the beam settings are copied. keepState:
 * the beam stays resident for pbs_server_submit_weights, under the handle *ticket. */
int pbs_server_submit(pbs_server *s, int ctId, const int doseDims[3], const float doseRes[3],
                      const float doseCorner[3], const pbs_beam_settings *beam, int keepState, long long *ticket);
/* Queue a weight-only update of a resident beam (numSpots weights in the beam's spot order). */
int pbs_server_submit_weights(pbs_server *s, long long stateTicket, const float *spotWeights, int numSpots,
                              long long *ticket);
/* 1 done, 0 queued or running, < 0 the request's error (pbs_last_error has the message). */
int pbs_server_query(pbs_server *s, long long ticket);
int pbs_server_wait(pbs_server *s, long long ticket);
/* Waits, copies the dose ([nz][ny][nx] float32) to host and frees the request's dose buffer. */
int pbs_server_result(pbs_server *s, long long ticket, float *doseHost);
/* The same into a device buffer of the server's GPU (no host transfer). */
int pbs_server_result_device(pbs_server *s, long long ticket, float *doseDevice);
/* Milliseconds from submit to admission and from submit to completion (-1 while pending). */
int pbs_server_timing(pbs_server *s, long long ticket, double *queueMs, double *totalMs);
/* Frees a resident beam (no update of it may be queued or running). */
int pbs_server_release_state(pbs_server *s, long long stateTicket);
int pbs_server_memory(pbs_server *s, double *usedBytes, double *budgetBytes);

#ifdef __cplusplus
}
#endif
#endif /* PBS_ABI_H */
