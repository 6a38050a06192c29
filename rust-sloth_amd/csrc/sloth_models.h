// Model-file ingestion (internal): OBJ and STL files to triangle soups the way
// the reference's loaders produce its SimpleMesh list.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace sloth {

struct SoupMesh {
    std::string name;
    std::vector<float> xyz;     // [n][3][3]
    std::vector<uint8_t> rgb;   // [n][3]
    float bbmin[3], bbmax[3];
    uint32_t n() const { return (uint32_t)(rgb.size() / 3); }
};

// Parses `path` by its extension (".obj" / ".stl", any case) into meshes in
// the reference's mesh-queue order.  Returns an empty string on success, else
// the error message.
std::string load_model_file(const std::string& path, std::vector<SoupMesh>& out);

}  // namespace sloth
