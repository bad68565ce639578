// zmpm_bgeo.hpp -- particle output as Houdini binary geometry (BGEO version 5).
//
// The reference writes its per-frame particle files through Partio v1.1.0
// (Partio::write on a "*.bgeo" name, from writePartio, Lib/Ziran/Math/
// Geometry/PartioIO.h:166-200, called by MpmSimulationBase::writeState,
// MpmSimulationBase.cpp:1066-1079, and by the AnisoMPM visualizer,
// Projects/anisofracture/AnisoVisualizer.h:14-63).  Partio is not linked
// here; this writer emits the point-only layout Partio's BGEO writer produces:
//
//   all integers and floats big-endian
//   int32 magic ('B' 'g' 'e' 'o'), char 'V', int32 version 5
//   int32 points, primitives (0), point groups (0), primitive groups (0)
//   int32 point attributes besides position, vertex attributes (0),
//         primitive attributes (0), detail attributes (0)
//   per point attribute: int16 name length, the name's bytes, int16 component
//         count, int32 type (5: Partio::VECTOR, which is what the reference
//         declares, also for one component), one int32 default 0 per component
//   per point: float x, y, z, w = 1, then each attribute's components as float
//   the extra-data markers: one byte 0x00, one byte 0xff
//
// Host code only.
#pragma once

#include <cstdint>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace zmpm {

// One per-point attribute: `count` components per point, point-major.
struct BgeoAttribute {
    std::string name;
    int count = 1;
    const double* data = nullptr; // [points][count]
};

namespace bgeo_detail {
inline void put32(std::vector<unsigned char>& b, uint32_t v)
{
    b.push_back((unsigned char)(v >> 24));
    b.push_back((unsigned char)(v >> 16));
    b.push_back((unsigned char)(v >> 8));
    b.push_back((unsigned char)v);
}
inline void put16(std::vector<unsigned char>& b, uint16_t v)
{
    b.push_back((unsigned char)(v >> 8));
    b.push_back((unsigned char)v);
}
inline void putFloat(std::vector<unsigned char>& b, double v)
{
    const float f = (float)v;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    put32(b, u);
}
} // namespace bgeo_detail

// Writes n points with positions x[n][3] and the given attributes to `path`.
inline void writeBgeoPoints(const std::string& path, int64_t n, const double* x, const std::vector<BgeoAttribute>& attrs)
{
    using namespace bgeo_detail;
    if (n < 0 || n > INT32_MAX || (n > 0 && !x)) throw std::invalid_argument("writeBgeoPoints: bad point array");
    for (const BgeoAttribute& a : attrs)
        if (a.name.empty() || a.name.size() > 32767 || a.name == "position" || a.count < 1 || a.count > 32767
            || (n > 0 && !a.data))
            throw std::invalid_argument("writeBgeoPoints: bad attribute " + a.name);
    std::vector<unsigned char> b;
    const uint32_t magic = ((((('B' << 8) | 'g') << 8) | 'e') << 8) | 'o';
    put32(b, magic);
    b.push_back('V');
    put32(b, 5);
    put32(b, (uint32_t)n);
    put32(b, 0); // primitives
    put32(b, 0); // point groups
    put32(b, 0); // primitive groups
    put32(b, (uint32_t)attrs.size());
    put32(b, 0); // vertex attributes
    put32(b, 0); // primitive attributes
    put32(b, 0); // detail attributes
    for (const BgeoAttribute& a : attrs) {
        put16(b, (uint16_t)a.name.size());
        b.insert(b.end(), a.name.begin(), a.name.end());
        put16(b, (uint16_t)a.count);
        put32(b, 5); // Partio::VECTOR
        for (int c = 0; c < a.count; c++) put32(b, 0);
    }
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("writeBgeoPoints: cannot open " + path);
    f.write((const char*)b.data(), (std::streamsize)b.size());
    std::vector<unsigned char> rec;
    for (int64_t p = 0; p < n; p++) {
        rec.clear();
        for (int c = 0; c < 3; c++) putFloat(rec, x[3 * p + c]);
        putFloat(rec, 1.0);
        for (const BgeoAttribute& a : attrs)
            for (int c = 0; c < a.count; c++) putFloat(rec, a.data[(size_t)p * a.count + c]);
        f.write((const char*)rec.data(), (std::streamsize)rec.size());
    }
    const unsigned char tail[2] = { 0x00, 0xff };
    f.write((const char*)tail, 2);
    if (!f) throw std::runtime_error("writeBgeoPoints: write failed: " + path);
}

} // namespace zmpm
