// A C++ caller of the drop-in subsecondWrapper (include/raytracedicom_integration.h, the
// reference's include/core/raytracedicom_integration.h:132-144): it fills RTDEnergyStruct and
// RTDBeamSettings (std::vector members) from a record file, calls subsecondWrapper on a host
// dose buffer and writes the dose back. tests/test_gpu_cpp_wrapper.py writes the records,
// runs this program and checks the dose against the FP64 oracle.
//
//   rtd_wrapper_main <records.bin> <dose_out.f32>     run subsecondWrapper
//   rtd_wrapper_main --parse <records.bin>            parse and print the struct sizes only
//
// Record file: "PBSR" then records of (u32 name length, name, u32 kind: 0 float32 / 1 int32,
// u64 count, data). Beam fields are named "b<index>.<field>" after the RTDBeamSettings members;
// energy fields "e.<member>"; the grid "ct", "ctDims", "ctRes", "ctCorner", "doseDims",
// "doseRes", "doseCorner", "doseIn", "nuclear".
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "raytracedicom_integration.h"

namespace {

struct Rec {
    uint32_t kind = 0;
    std::vector<float> f;
    std::vector<int> i;
};

bool read_records(const char *path, std::map<std::string, Rec> &out) {
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return false;
    char magic[4];
    bool ok = std::fread(magic, 1, 4, fp) == 4 && std::memcmp(magic, "PBSR", 4) == 0;
    while (ok) {
        uint32_t n = 0;
        if (std::fread(&n, 4, 1, fp) != 1) break;  // end of file
        std::string name(n, '\0');
        Rec r;
        uint64_t count = 0;
        ok = std::fread(&name[0], 1, n, fp) == n && std::fread(&r.kind, 4, 1, fp) == 1 &&
             std::fread(&count, 8, 1, fp) == 1 && r.kind <= 1;
        if (!ok) break;
        if (r.kind == 0) {
            r.f.resize(count);
            ok = std::fread(r.f.data(), 4, count, fp) == count;
        } else {
            r.i.resize(count);
            ok = std::fread(r.i.data(), 4, count, fp) == count;
        }
        out[name] = std::move(r);
    }
    std::fclose(fp);
    return ok;
}

struct Reader {
    const std::map<std::string, Rec> &m;
    bool has(const std::string &k) const { return m.count(k) != 0; }
    std::vector<float> f(const std::string &k) const {
        auto it = m.find(k);
        return it == m.end() ? std::vector<float>() : it->second.f;
    }
    std::vector<int> i(const std::string &k) const {
        auto it = m.find(k);
        return it == m.end() ? std::vector<int>() : it->second.i;
    }
    float f1(const std::string &k, float d = 0.0f) const {
        const std::vector<float> v = f(k);
        return v.empty() ? d : v[0];
    }
    int i1(const std::string &k, int d = 0) const {
        const std::vector<int> v = i(k);
        return v.empty() ? d : v[0];
    }
    float2 f2(const std::string &k) const {
        const std::vector<float> v = f(k);
        return v.size() >= 2 ? float2{v[0], v[1]} : float2{0.0f, 0.0f};
    }
    float3 f3(const std::string &k, float3 d) const {
        const std::vector<float> v = f(k);
        return v.size() >= 3 ? float3{v[0], v[1], v[2]} : d;
    }
    int3 i3(const std::string &k) const {
        const std::vector<int> v = i(k);
        return v.size() >= 3 ? int3{v[0], v[1], v[2]} : int3{0, 0, 0};
    }
};

RTDBeamSettings beam_from(const Reader &r, int b) {
    const std::string p = "b" + std::to_string(b) + ".";
    RTDBeamSettings s;
    s.energies = r.f(p + "energies");
    const std::vector<float> sig = r.f(p + "spotSigmas");
    for (size_t l = 0; l + 1 < sig.size(); l += 2) s.spotSigmas.push_back(float2{sig[l], sig[l + 1]});
    s.raySpacing = r.f2(p + "raySpacing");
    s.steps = unsigned(r.i1(p + "steps"));
    s.sourceDist = r.f2(p + "sourceDist");
    s.spotOffset = r.f3(p + "spotOffset", float3{0.0f, 0.0f, 0.0f});
    s.spotDelta = r.f3(p + "spotDelta", float3{0.0f, 0.0f, 0.0f});
    s.maxSubspotsPerLayer = r.i1(p + "maxSubspotsPerLayer");
    s.subspotData = r.f(p + "subspotData");
    s.layerSpotCounts = r.i(p + "layerSpotCounts");
    s.spotPositions = r.f(p + "spotPositions");
    s.spotPositionsAreIndices = r.i1(p + "spotPositionsAreIndices") != 0;
    s.spotWeights = r.f(p + "spotWeights");
    s.spotBeamDirections = r.f(p + "spotBeamDirections");
    s.layerLongitudinalCutoffs = r.f(p + "layerLongitudinalCutoffs");
    s.waterEquivalence = r.f(p + "waterEquivalence");
    s.beamDirection = r.f3(p + "beamDirection", s.beamDirection);
    s.bmXDirection = r.f3(p + "bmXDirection", s.bmXDirection);
    s.bmYDirection = r.f3(p + "bmYDirection", s.bmYDirection);
    s.sourcePosition = r.f3(p + "sourcePosition", s.sourcePosition);
    s.sad = r.f1(p + "sad");
    s.refPlaneZ = r.f1(p + "refPlaneZ");
    s.roiLinearIndices = r.i(p + "roiLinearIndices");
    return s;
}

RTDEnergyStruct energy_from(const Reader &r) {
    RTDEnergyStruct e;
    e.nEnergySamples = r.i1("e.nEnergySamples");
    e.nEnergies = r.i1("e.nEnergies");
    e.energiesPerU = r.f("e.energiesPerU");
    e.peakDepths = r.f("e.peakDepths");
    e.scaleFacts = r.f("e.scaleFacts");
    e.ciddMatrix = r.f("e.ciddMatrix");
    e.nDensitySamples = r.i1("e.nDensitySamples");
    e.densityScaleFact = r.f1("e.densityScaleFact");
    e.densityVector = r.f("e.densityVector");
    e.nSpSamples = r.i1("e.nSpSamples");
    e.spScaleFact = r.f1("e.spScaleFact");
    e.spVector = r.f("e.spVector");
    e.nRRlSamples = r.i1("e.nRRlSamples");
    e.rRlScaleFact = r.f1("e.rRlScaleFact");
    e.rRlVector = r.f("e.rRlVector");
    e.nucWeightMatrix = r.f("e.nucWeightMatrix");
    e.nucSqSigmaMatrix = r.f("e.nucSqSigmaMatrix");
    return e;
}

}  // namespace

int main(int argc, char **argv) {
    const bool parseOnly = argc == 3 && std::strcmp(argv[1], "--parse") == 0;
    if (!parseOnly && argc != 3) {
        std::fprintf(stderr, "usage: %s <records.bin> <dose_out.f32> | --parse <records.bin>\n", argv[0]);
        return 2;
    }
    std::map<std::string, Rec> recs;
    if (!read_records(parseOnly ? argv[2] : argv[1], recs)) {
        std::fprintf(stderr, "cannot read the record file\n");
        return 2;
    }
    const Reader r{recs};
    const int numBeams = r.i1("numBeams");
    std::vector<RTDBeamSettings> beams;
    for (int b = 0; b < numBeams; ++b) beams.push_back(beam_from(r, b));
    const RTDEnergyStruct energy = energy_from(r);
    const std::vector<float> ct = r.f("ct");
    const int3 ctDims = r.i3("ctDims"), doseDims = r.i3("doseDims");
    const float3 ctRes = r.f3("ctRes", float3{0, 0, 0}), ctCorner = r.f3("ctCorner", float3{0, 0, 0});
    const float3 doseRes = r.f3("doseRes", float3{0, 0, 0}), doseCorner = r.f3("doseCorner", float3{0, 0, 0});
    std::vector<float> dose = r.f("doseIn");
    const size_t nd = size_t(doseDims.x) * doseDims.y * doseDims.z;
    if (dose.size() != nd) dose.assign(nd, 0.0f);
    if (parseOnly) {
        std::printf("beams %d layers %zu subspots %zu ct %zu dose %zu energies %d\n", numBeams,
                    beams.empty() ? size_t(0) : beams[0].energies.size(),
                    beams.empty() ? size_t(0) : beams[0].subspotData.size(), ct.size(), dose.size(), energy.nEnergies);
        return 0;
    }
    subsecondWrapper(ct.data(), ctDims, ctRes, ctCorner, dose.data(), doseDims, doseRes, doseCorner, beams.data(),
                     beams.size(), &energy, 0, r.i1("nuclear") != 0, 0);
    FILE *fp = std::fopen(argv[2], "wb");
    if (!fp || std::fwrite(dose.data(), 4, nd, fp) != nd) {
        std::fprintf(stderr, "cannot write the dose\n");
        if (fp) std::fclose(fp);
        return 2;
    }
    std::fclose(fp);
    return 0;
}
