/*
 * FP64 CPU oracle of the pencil-beam final-dose path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP engine: it is loaded
 * by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by nothing
 * that ships. The engine never calls into it.
 *
 * It restates, stage by stage, the method of subsecondWrapper
 * (src/core/raytracedicom_wrapper.cu:2446-5661) and the kernels it launches. Host
 * planning (units, beam basis, spot lattice, CPB/ray grid, step counts, per-layer LUT
 * parameters) is evaluated in float exactly as the reference host code does, because the
 * integers it produces (grid sizes, corners, step ranges) are defined by that float
 * arithmetic. Every per-ray, per-step, per-pixel and per-voxel quantity (trilinear CT
 * samples, WEPL, convolution taps, IDD, sigma, superposition, BEV->dose transfer) is
 * evaluated in double. bench.py times this library on the host cores as the CPU baseline.
 *
 * Parity pinning (which check pins which function) is listed next to each declaration.
 */
#ifndef PBS_ORACLE_H
#define PBS_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_LAYERS 256
#define ORC_MAX_RAD 32
#define ORC_TILE_X 32
#define ORC_TILE_Y 8
#define ORC_MIN_TILES_IN_BATCH 16

typedef struct {
    int nEnergies, nSamples;
    const float *energies, *peaks, *scales, *cidd; /* cidd [nEnergies][nSamples] */
    int nDensity; float densityScale; const float *density;
    int nSp; float spScale; const float *sp;
    int nRrl; float rrlScale; const float *rrl;
    const float *nucWeight, *nucSqSigma; /* optional [nEnergies][nSamples] */
    /* ion of the table set: residual energy c (peak - depth)^q MeV/u, rest mass per nucleon,
     * (Z/A)^2 on the scattering constant; 0 = proton (idd_sigma.cu:172-181) */
    float ionRangeCoef, ionRangeExp, ionRestMass, ionScatterScale;
} orc_energy;

typedef struct {
    int numLayers;
    const float *energies;    /* [L] MeV/u */
    const float *spotSigmas;  /* [L][2] at iso in air */
    float raySpacing[2];
    float sourceDist[2];
    float spotDelta[3];
    int maxSubspots;
    const float *subspotData; /* [L][maxSubspots][5] = dx, dy, w, sx, sy */
    const int *layerSpotCounts; /* [L], NULL -> CPB (legacy) path */
    int numSpots;
    const float *spotPositions; /* [numSpots][2] gantry-frame x,y on the iso plane */
    const float *spotWeights;   /* [numSpots] */
    const float *spotBeamDirections; /* [numSpots][3] or NULL */
    float beamDirection[3], bmXDirection[3], bmYDirection[3], sourcePosition[3];
    float sad;
    int useSubspotConvolution; /* 0 = per-layer spot sigma (reference default) */
    int nuclear;               /* halo branch: 0 off, 1 GAUSS_FIT, 2 SOUKUP, 3 FLUKA constants */
    /* WEQ input (raytracedicom_integration.h:63-70): 9-value header [depth_start, depth_step,
     * depth_n, y_start, y_step, y_n, x_start, x_step, x_n] + samples [x][y][k], or NULL */
    const float *weq;
    size_t weqCount;
    int spotIndices;           /* spot positions are WEQ-lattice texel coordinates */
    /* explicit CPB grid in the gantry frame (cudaFinalDose: the ROI box at 0.5 mm,
     * cuda_final_dose.cu:302-308); 0 = the wrapper's own grid */
    int cpbOverride;
    float cpbOvCorner[2], cpbOvRes[2];
    int cpbOvDims[2];
    /* the environment overrides of the units (wrapper:2742-2790, 3193-3202), as values:
     * energyDepthMode 0 = the cm-CT / mm-table heuristic, 1 = energyDepthToMm as given,
     * 2 = the CT's length unit; cpbCoverage > 0 replaces the legacy CPB coverage */
    int energyDepthMode;
    float energyDepthToMm;
    float cpbCoverage;
    /* CarbonPBS lateral-profile model and its hooks (bit 0 sigma, bit 1 overall weight; the
     * reference has both off, wrapper:2812-2814) */
    const float *profileEnergies;
    int numProfileEnergies;
    const float *profileData;   /* [rows][depthN][channels] */
    size_t profileDataCount;
    const float *profileSetting;  /* depth0, depthStep, depthN */
    int numProfileSetting;
    const float *beamParaData;  /* [rows][3] */
    int numBeamParaData;
    float beamParaPos;
    int carbonProfileHooks;
} orc_beam;

/* Beam plan: everything the host computes before and between kernels. */
typedef struct {
    int ok;                 /* 1 on success */
    int legacyPath;         /* 1 = subspot->CPB path (no spot lattice) */
    float lenToMm, energyDepthToMm;
    float bmX[3], bmY[3], bmZ[3], iso[3];
    float g2w[9], g2wOff[3];      /* gantry -> world (row-major) */
    float g2im[9], g2imOff[3];    /* gantry -> CT index */
    float g2dose[9], g2doseOff[3];/* gantry -> dose index */
    float sourceDist[2];
    /* spot lattice */
    int spotDims[2];
    float spotOffset[2], spotDelta[2];
    /* CPB / ray grid */
    float cpbCorner[3], cpbRes[3];
    int rayDims[2];         /* rays traced (lattice path: multiples of 32x8) */
    int superpDims[2];      /* rays padded to 32x8 for the superposition */
    int steps;              /* tracer steps K */
    float fanCorner[3], fanDelta[3];
    /* per-beam trace results */
    int beamFirstInside, beamFirstOutside, firstPastCutoffAll, guaranteedPassive;
    float maxPeakDepth;
    /* per-layer host parameters */
    float energyIdx[ORC_MAX_LAYERS], scaleFact[ORC_MAX_LAYERS], peakDepth[ORC_MAX_LAYERS];
    float convPeakDepth[ORC_MAX_LAYERS];
    float entrySigma[ORC_MAX_LAYERS][2];
    float sigmaSqAirLin[ORC_MAX_LAYERS], sigmaSqAirQuad[ORC_MAX_LAYERS];
    float volConst, volLin, volSq;
    float spotDistInRays;
    int afterLast[ORC_MAX_LAYERS];
    int layerPassive[ORC_MAX_LAYERS];
    int radHist[ORC_MAX_LAYERS][ORC_MAX_RAD + 2];  /* tiles per radius bin (bin 33 = overflow) */
    int batchRad[ORC_MAX_LAYERS][ORC_MAX_RAD + 1]; /* tiles per launch radius after batching */
    int transferStart[3], transferMax[3];         /* union box over layers */
    int numLayers;
    char error[256];
    /* halo grid (nuclear branch): spot lattice rounded up to 32x8, spot spacing in rays */
    int nucDims[2];
    float spotDist;
    int nucRadHist[ORC_MAX_LAYERS][ORC_MAX_RAD + 2];
    float entrySpotSigmaX[ORC_MAX_LAYERS];  /* spot sigma_x in mm (halo entry sigma) */
    /* WEQ-volume input: samples replace the CT trace (wrapper:1923-1988, 3305-3370) */
    int weqVolume, spotIndices;
    int weqDims[3];                          /* x_n, y_n, depth_n */
    float weqStepMm, weqX0, weqY0, weqDx, weqDy;
    const float *weqBody;                    /* [x_n][y_n][depth_n] */
    /* halo lattice (cells = spots): the spot lattice for physical positions, the physical PB
     * view of index-coded spots otherwise (wrapper:1478-1820, 3421-3437) */
    int nucRaw[2], nucFallback;
    float nucOffset[2], nucDelta[2];
    /* profile hooks that apply (wrapper:2816-2862, 4577-4603): table shape and depth axis, per
     * layer the fractional table row and the beam parameters interpolated there */
    int profileHooks, profileRows, profileDepthN, profileChannels;
    float profileDepth0, profileDepthStep, beamParaPos;
    float profileLayer[ORC_MAX_LAYERS][4];
    const float *profileTable;
} orc_plan;

long orc_sizeof(int which);  /* 0 orc_beam, 1 orc_plan, 2 orc_energy */
/* OpenMP threads of the following calls (the CPU baseline's 1-thread and all-core timings). */
void orc_set_threads(int n);

/* ---- helpers exposed for known-answer tests ---- */
/* findDecimalOrdered (include/utils/vector_find.h:130-145). Pinned: KAT peak depths. */
float orc_find_decimal_ordered(const float *v, int n, float value);
/* vectorInterpolate (include/utils/vector_interpolate.h:19-30). Pinned: KAT peak depths. */
float orc_vector_interpolate(const float *v, int n, float idx);
/* findFirstLargerOrdered (include/utils/vector_find.h:62-81). Pinned: brute force. */
int orc_find_first_larger_ordered(const double *v, int n, double value);
/* alignGridToReferencePhase (wrapper:970-997). Pinned: hand-computed config grids. */
int orc_align_grid(float minC, float maxC, float ref, float step, int multiple,
                   float *origin, int *count, int *first, int *last);
/* Per-tile radius (complete_superposition.cu:242): float formula. Pinned: brute force. */
int orc_tile_radius_f32(float minRSigma);
/* Batching of radius bins into launches (complete_superposition.cu:465-490). */
void orc_batch_bins(const int *hist /*[34]*/, int *batched /*[33]*/, int *effRad /*[34]*/);
/* Tile radii for a [K][ry][rx] float field over steps [k0,k1): rad[(k-k0)][ty][tx].
 * NaN entries are ignored (fminf semantics). Pinned: brute-force min + integer formula. */
void orc_tile_radii_f32(const float *rsigma, int rx, int ry, int k0, int k1, int *rad);

/* CT adapter (wrapper:2307-2441): sampled min/max decide the input type unless forcedType >= 0
 * (0 HU, 1 HU+1000, 2 density, 3 SPR; the RTD_CT_INPUT_TYPE override); HU gets +1000, density and
 * SPR are inverted through the density / SP table by findFractionalIndexMonotonic's bisection
 * (wrapper:2337-2375). out[n]. Returns the type used, or -1 when the table is empty.
 * Pinned: HU shift; table round trip (the table at the result equals the input) on every value
 * class and the bisection's crossing among all crossings on the non-monotonic density stretch. */
int orc_ct_adapter(const float *in, size_t n, int forcedType, const orc_energy *en, float *out);

/* ---- stages ---- */
/* Host planning up to (but excluding) the trace. Pinned: closed-form grids of the configs. */
int orc_plan_beam(const int ctDims[3], const float ctRes[3], const float ctCorner[3],
                  const int doseDims[3], const float doseRes[3], const float doseCorner[3],
                  const orc_beam *beam, const orc_energy *en, orc_plan *plan);

/* BEV ray trace (bev_ray_tracing.cu:65-159). Outputs [K][ry][rx] and [ry][rx].
 * Pinned: homogeneous water WEPL(k) = (k+1) stepLen, slab WEPL, voxel-centre samples. */
void orc_trace(const orc_plan *plan, const float *ctHuPlus1000, const int ctDims[3],
               const orc_energy *en, double *density, double *wepl, int *firstInside, int *firstOutside);

/* Distance of each ray's entry decision from the threshold (bev_ray_tracing.cu:129-141): with
 * c(k) the cumulative HU+1000 after step k (non-decreasing) and f = firstInside, the smaller of
 * 150 - c(f-1) (when c(f-1) < 150) and c(f) - 150 (when f < K); +inf when neither exists or for a
 * WEQ volume. A ray whose margin is below the trace's sample rounding can enter one step earlier
 * or later in a float trace. [ry][rx]. Pinned: a slab CT whose cumulative HU lands on 150 +- e. */
void orc_trace_entry_margin(const orc_plan *plan, const float *ctHuPlus1000, const int ctDims[3],
                            double *margin);

/* Per-beam reductions and cutoffs (wrapper:3651-3700). Fills plan->beamFirstInside.. */
void orc_beam_cutoffs(orc_plan *plan, const orc_beam *beam, const orc_energy *en,
                      const double *wepl, const int *firstInside, const int *firstOutside,
                      double *weplMin /*[K]*/);

/* Spot-lattice -> ray weights, all layers (wrapper:3733-3875, convolution.cu:198-262, 477-513).
 * Pinned: telescoping sum of taps = 1/2[erf(hi)-erf(lo)], brute-force 2D Gaussian integral. */
void orc_conv_lattice(orc_plan *plan, const orc_beam *beam, const orc_energy *en, double *rayWeights);

/* Subspot -> CPB 2D erf convolution for one layer-stack (convolution.cu:21-152, 342-450) and
 * the area-overlap CPB->ray map (convolution.cu:269-339). Used by the CPB path. */
void orc_subspot_to_cpb(const float *subspotData, int numLayers, int maxSubspots,
                        const float cpbCorner[3], const float cpbRes[3], const int cpbDims[3],
                        const float beamDir[3], const float bmX[3], const float bmY[3],
                        const float source[3], float sad, double *cpb);
void orc_cpb_to_rays(const double *cpb, const int cpbDims[3], const float cpbCorner[3], const float cpbRes[3],
                     double *rays, const int rayDims[2], const float rayCorner[3], const float rayRes[3], int layer);

/* Per-layer parameters (wrapper:3917-4386). Requires cutoffs. */
void orc_layer_params(orc_plan *plan, const orc_beam *beam, const orc_energy *en, const double *weplMin);

/* IDD and sigma fill for one layer (idd_sigma.cu:46-259). Outputs [K][ry][rx] over
 * [first, afterLast); other steps are left untouched. Pinned: telescoping energy sum,
 * independent recurrence evaluation. */
void orc_fill(const orc_plan *plan, int layer, const orc_energy *en, const double *density,
              const double *wepl, const double *rayWeights, const int *firstInside,
              const int *firstOutside, double *idd, double *rsigma, int *firstPassive);

/* CarbonPBS profile hooks on layer l's field after the fill (overrideRSigmaFromCarbonProfileKernel,
 * applyCarbonProfileOverallWeightKernel; wrapper:2080-2188, 4604-4832), steps [beamFirstInside,
 * afterLast) of the ray grid ([K][ry][rx]): sigma from the profile model where IDD > 0 (+inf
 * elsewhere), then IDD times the overall-weight channel. The table is read by exact bilinear
 * interpolation over depth and row (the reference's texture unit uses 8-bit weights: parity
 * unpinned there). Pinned: closed forms of a one-Gaussian table (sigma override, divergence
 * width, physical-depth polynomial) and of a depth-linear overall weight. */
void orc_profile_hooks(const orc_plan *plan, int layer, const double *wepl, double *idd, double *rsigma);

/* Halo lattice plan (wrapper:804-929; spot spacing wrapper:4062-4100, 4149-4172, 4338-4350):
 * plan->nucDims = spot lattice dims rounded up to 32x8, plan->spotDist = spot spacing / fan
 * ray spacing; rayToNuc[ray] = halo cell of the spot whose gantry position rounds to that ray
 * (else -1); nucWeights[L][nucY][nucX] = the layer lattice weights (may be NULL).
 * Returns 0, or -1 with plan->error (a spot outside the ray grid, spotDist < 1.5, no
 * lattice). Pinned: hand-computed config-1 plan (11x11 spots at 5 rays). */
int orc_halo_plan(orc_plan *plan, const orc_beam *beam, int *rayToNuc, double *nucWeights);
/* IDD / sigma fill with the halo branch (idd_sigma.cu:46-259 under NUCLEAR_CORR): model 0 is
 * exactly orc_fill; otherwise the primary takes (1 - w_nuc) of the energy and the model's
 * scattering constants, and each spot's central ray writes the halo IDD and rSigmaEff into
 * nucIdd / nucRsigma [K][nucY][nucX] (nucW = this layer's halo weights [nucY][nucX]).
 * Pinned: energy split sum_k (res + nucRes spotDist^2) mass = w CIDD telescoping. */
void orc_fill_halo(const orc_plan *plan, int layer, const orc_energy *en, const double *density,
                   const double *wepl, const double *rayWeights, const int *firstInside,
                   const int *firstOutside, double *idd, double *rsigma, int *firstPassive, int model,
                   const int *rayToNuc, const double *nucW, double *nucIdd, double *nucRsigma);
/* Superposition of one layer into a BEV volume [bevZ][ry+64][rx+64]
 * (complete_superposition.cu:197-544). Also records radius histogram and batching.
 * Pinned: per-pixel kernel mass erf(r(R+1/2))^2, brute-force scatter. */
void orc_superpose(orc_plan *plan, int layer, const double *idd, const double *rsigma, double *bev);

/* BEV -> dose trilinear gather for one layer (wrapper:5343-5505, prim_transf_kernel.cu:11-44).
 * Pinned: identity geometry, linearity. */
void orc_transfer(orc_plan *plan, int layer, const double *bev, double *dose,
                  const int doseDims[3]);

/* Whole beam, all layers: accumulates into dose (double [nz][ny][nx]). Returns 0 on success. */
int orc_beam_dose(const float *ctHuPlus1000, const int ctDims[3], const float ctRes[3], const float ctCorner[3],
                  double *dose, const int doseDims[3], const float doseRes[3], const float doseCorner[3],
                  const orc_beam *beam, const orc_energy *en, orc_plan *planOut);

/* orc_beam_dose with the trace's firstInside replaced by entryOverride[r] wherever that is >= 0
 * ([ry][rx]; NULL: none). Lets a test follow a float trace's entry step on rays whose decision
 * sits on the threshold (orc_trace_entry_margin). */
int orc_beam_dose_entry(const float *ctHuPlus1000, const int ctDims[3], const float ctRes[3],
                        const float ctCorner[3], double *dose, const int doseDims[3], const float doseRes[3],
                        const float doseCorner[3], const orc_beam *beam, const orc_energy *en,
                        const int *entryOverride, orc_plan *planOut);

/* Trilinear sample with border 0 at continuous index (x,y,z), voxel centres at integers
 * (bev_ray_tracing.cu:17-62 after its -0.5). Exposed for tests. */
double orc_sample3d_border(const float *vol, const int dims[3], double x, double y, double z);

/* Adjoint with respect to the spot weights (spot-lattice path, halo off): gSpots[numSpots] =
 * dJ/dw for J with dJ/dDose = gDose ([nz][ny][nx]), at the weights of `beam`. The transposes of
 * the transfer, superposition, IDD scaling and separable spot convolution, in FP64.
 * Pinned: <gDose, dose(w)> = <w, gSpots> (the dose is linear in w for fixed masks) and central
 * finite differences of orc_beam_dose. */
int orc_beam_adjoint(const float *ctHuPlus1000, const int ctDims[3], const float ctRes[3], const float ctCorner[3],
                     const double *gDose, const int doseDims[3], const float doseRes[3], const float doseCorner[3],
                     const orc_beam *beam, const orc_energy *en, double *gSpots, orc_plan *plan);
#ifdef __cplusplus
}
#endif
#endif
