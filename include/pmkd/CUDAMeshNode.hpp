// CUDAMeshNode.hpp -- an indexed mesh in device memory
// (Scene/CUDAMeshNode.h:20-41): float4 vertices (w = 1), float4 normals,
// uchar4 colors, uint indices; GetSize() is the index count.  Built from
// host arrays instead of an OpenEngine MeshNode (CUDAMeshNode.cu:72-139).
#pragma once
#include <vector>

#include "DeviceBuffer.hpp"

namespace pmkd {
namespace Scene {

class CUDAMeshNode {
public:
    CUDAMeshNode(const std::vector<float4> &vertices, const std::vector<unsigned> &indices,
                 const std::vector<float4> &normals = {}, const std::vector<uchar4> &colors = {}) {
        if (indices.size() % 3) throw Exception(PM_ERR_ARG, "index count is not a multiple of 3");
        vertices_.Upload(vertices);
        indices_.Upload(indices);
        normals_.Upload(normals.empty() ? std::vector<float4>(vertices.size(), float4{0, 0, 0, 0}) : normals);
        // CUDAMeshNode.cu:133: default color (180, 180, 180, 255)
        colors_.Upload(colors.empty() ? std::vector<uchar4>(vertices.size(), uchar4{180, 180, 180, 255}) : colors);
    }
    float4 *GetVertexData() const { return vertices_.GetDeviceData(); }
    float4 *GetNormalData() const { return normals_.GetDeviceData(); }
    uchar4 *GetColorData() const { return colors_.GetDeviceData(); }
    unsigned int *GetIndexData() const { return indices_.GetDeviceData(); }
    unsigned int GetSize() const { return (unsigned int)indices_.GetSize(); }
    int GetVertexCount() const { return (int)vertices_.GetSize(); }

private:
    DeviceBuffer<float4> vertices_, normals_;
    DeviceBuffer<uchar4> colors_;
    DeviceBuffer<unsigned> indices_;
};

}  // namespace Scene
}  // namespace pmkd
