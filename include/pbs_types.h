/*
 * Plain-data types of the pencil-beam dose path, laid out byte-for-byte like the
 * reference's public types so that buffers built for it can be handed over unchanged.
 *
 *   vec3f, vec3i, vec2f ....... include/core/common.cuh:39-78
 *   Grid (48 B) ............... include/core/common.cuh:146-149
 *   Source (40 B) ............. include/core/common.cuh:151-154
 *   Beam (52 B) ............... include/core/common.cuh:156-160
 *   SubspotInfo (56 B) ........ include/core/common.cuh:163-217 (data members only)
 *   CPBGrid (40 B) ............ include/core/common.cuh:220-228 (data members only)
 *
 * The reference gives these types constructors and member functions; here they are
 * aggregates usable from C. Member order, types and padding are identical.
 */
#ifndef PBS_TYPES_H
#define PBS_TYPES_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vec3f { float x, y, z; } vec3f;
typedef struct vec3i { int x, y, z; } vec3i;
typedef struct vec2f { float x, y; } vec2f;

typedef struct Grid {
    vec3f corner, resolution, upperCorner;
    vec3i dims;
} Grid;

typedef struct Source {
    vec3f sourcePos, isoPos, vsDirX;
    float vSAD;
} Source;

typedef struct Beam {
    vec3f source, bmdir, bmxdir;
    float ene, transCutoff, longitudalCutoff;
    int eneIdx;
} Beam;

/* One subspot of a layer: offsets on the reference plane, weight, sigmas; the derived
 * position/direction/effective sigma are filled by the engine (common.cuh:172-217). */
typedef struct SubspotInfo {
    float deltaX, deltaY, weight, sigmaX, sigmaY;
    int eneIdx;
    vec3f position, direction;
    float effectiveSigma;
    unsigned char isValid; /* C++ bool in the reference: one byte, 3 bytes tail padding */
} SubspotInfo;

typedef struct CPBGrid {
    vec3f corner, resolution;
    vec3i dims;
    int nLayers;
} CPBGrid;

#ifdef __cplusplus
} /* extern "C" */
static_assert(sizeof(vec3f) == 12, "vec3f layout");
static_assert(sizeof(vec3i) == 12, "vec3i layout");
static_assert(sizeof(Grid) == 48, "Grid layout");
static_assert(sizeof(Source) == 40, "Source layout");
static_assert(sizeof(Beam) == 52, "Beam layout");
static_assert(sizeof(SubspotInfo) == 56, "SubspotInfo layout");
static_assert(sizeof(CPBGrid) == 40, "CPBGrid layout");
#else
_Static_assert(sizeof(Grid) == 48, "Grid layout");
_Static_assert(sizeof(Source) == 40, "Source layout");
_Static_assert(sizeof(Beam) == 52, "Beam layout");
_Static_assert(sizeof(SubspotInfo) == 56, "SubspotInfo layout");
_Static_assert(sizeof(CPBGrid) == 40, "CPBGrid layout");
#endif

#endif /* PBS_TYPES_H */
