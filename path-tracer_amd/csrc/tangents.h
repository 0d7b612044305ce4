// tangents.h - MikkTSpace tangent frames for the glTF front end.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace ptcore {

// One triangle list: index_count / 3 faces over per-vertex positions (xyz),
// unit normals (xyz) and texcoords (uv).
struct MeshView {
    const uint32_t* indices;
    size_t index_count;
    const float* positions;
    const float* normals;
    const float* texcoords;
};

// Tangent (xyz) and handedness sign (+1 orientation-preserving, -1 otherwise)
// per face corner, 4 floats each, face-major.
bool generate_tangents(const MeshView& mesh, std::vector<float>& out);

}  // namespace ptcore
