// Image readers/writers for the registration facade: raw files described by the YAML
// (Interpolator2DTo3D's path, interpolator.cu:1550-1627) and NRRD with raw encoding,
// which covers the reference's exports (interpolator.cu:681-920). The reference reads
// other formats through ITK (volumeImporter.cpp:42-93); those are not supported here.
#pragma once

#include <string>
#include <vector>

#include "lc2e.h"

namespace lc2io {

struct Image {
    std::vector<unsigned char> bytes;  // raw voxel data, x fastest
    lc2e_dtype dtype = LC2E_F32;
    lc2e_geom geom{};
};

// dtype names: float32, float16, int16, uint8 (also float, half, short, uchar).
lc2e_dtype parse_dtype(const std::string& name);
Image read_raw(const std::string& path, const int dims[3], const float spacing[3],
               const float origin[3], lc2e_dtype dtype);
Image read_nrrd(const std::string& path);
// Chooses by extension (.nrrd / .nhdr -> NRRD, anything else -> raw with the given geometry).
Image read_image(const std::string& path, const int dims[3], const float spacing[3],
                 const float origin[3], lc2e_dtype dtype);
void write_nrrd(const std::string& path, const float* data, const lc2e_geom& g);
void write_raw(const std::string& path, const void* data, size_t bytes);

}  // namespace lc2io
