// Wavefront OBJ mesh reader (host side) following e8util::wavefront_obj::load_geometry
// (resource.cpp:137-255).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace x3d {

struct ObjMesh {
    std::vector<float> verts;   // 3 per vertex, in file order
    std::vector<float> normals; // 3 per vertex: the normal of the last face corner naming it
    std::vector<float> uvs;     // 2 per vertex: likewise the texture coordinate
    std::vector<uint32_t> tris; // 3 per triangle, 0-based vertex indices
};

// Reads `path`; returns an empty string or an error message naming the line.
std::string load_obj(const char *path, ObjMesh &out);

} // namespace x3d
