// GeometryList.hpp -- triangle SoA p0/p1/p2 (float4) gathered from meshes
// (Utils/CUDA/GeometryList.h:36-89; AddMesh(CUDAMeshNode*, modelMat)
// GeometryList.cu:318-341) with the per-corner normals n0..n2 (float4) and
// colours c0..c2 (uchar4) the canvas shading reads.  Woop values are made
// by the tracers per call (GetWoopValues, GeometryList.cu:103-117).
#pragma once
#include <vector>

#include "CUDAMeshNode.hpp"

namespace pmkd {
namespace Utils {

class GeometryList {
public:
    GeometryList() { check(nullptr, pm_create(&ctx_, 0) == PM_OK ? PM_OK : PM_ERR_HIP); }
    ~GeometryList() { if (ctx_) pm_destroy(ctx_); }
    GeometryList(const GeometryList &) = delete;
    GeometryList &operator=(const GeometryList &) = delete;

    int GetSize() const { return size_; }
    float4 *GetP0Data() const { return p0_.GetDeviceData(); }
    float4 *GetP1Data() const { return p1_.GetDeviceData(); }
    float4 *GetP2Data() const { return p2_.GetDeviceData(); }
    float4 *GetNormal0Data() const { return n0_.GetDeviceData(); }
    float4 *GetNormal1Data() const { return n1_.GetDeviceData(); }
    float4 *GetNormal2Data() const { return n2_.GetDeviceData(); }
    uchar4 *GetColor0Data() const { return c0_.GetDeviceData(); }
    uchar4 *GetColor1Data() const { return c1_.GetDeviceData(); }
    uchar4 *GetColor2Data() const { return c2_.GetDeviceData(); }

    // modelRowMajor: p = M v (row i dotted with v); nullptr = identity
    void AddMesh(const Scene::CUDAMeshNode *mesh, const float *modelRowMajor = nullptr,
                 hipStream_t stream = nullptr) {
        const int tris = (int)mesh->GetSize() / 3;
        Extend(size_ + tris);
        check(ctx_, pm_add_mesh(ctx_, (const float *)mesh->GetVertexData(), mesh->GetVertexCount(),
                                mesh->GetIndexData(), (int)mesh->GetSize(), modelRowMajor,
                                (float *)p0_.GetDeviceData(), (float *)p1_.GetDeviceData(),
                                (float *)p2_.GetDeviceData(), size_, stream));
        check(ctx_, pm_add_mesh_attributes(ctx_, (const float *)mesh->GetNormalData(),
                                           (const unsigned char *)mesh->GetColorData(), mesh->GetVertexCount(),
                                           mesh->GetIndexData(), (int)mesh->GetSize(), modelRowMajor,
                                           (float *)n0_.GetDeviceData(), (float *)n1_.GetDeviceData(),
                                           (float *)n2_.GetDeviceData(), (unsigned char *)c0_.GetDeviceData(),
                                           (unsigned char *)c1_.GetDeviceData(),
                                           (unsigned char *)c2_.GetDeviceData(), size_, stream));
        size_ += tris;
    }
    // replace the contents with host triangles (3 float4 per triangle)
    void SetTriangles(const std::vector<float4> &p0, const std::vector<float4> &p1,
                      const std::vector<float4> &p2) {
        if (p0.size() != p1.size() || p0.size() != p2.size()) throw Exception(PM_ERR_ARG, "size mismatch");
        p0_.Upload(p0); p1_.Upload(p1); p2_.Upload(p2);
        size_ = (int)p0.size();
    }
    // CollectGeometry starts a new frame (GeometryList.cu:343-347)
    void Clear() { size_ = 0; }

private:
    void Extend(int n) {
        p0_.Extend(n); p1_.Extend(n); p2_.Extend(n);
        n0_.Extend(n); n1_.Extend(n); n2_.Extend(n);
        c0_.Extend(n); c1_.Extend(n); c2_.Extend(n);
    }
    pm_ctx *ctx_ = nullptr;
    DeviceBuffer<float4> p0_, p1_, p2_, n0_, n1_, n2_;
    DeviceBuffer<uchar4> c0_, c1_, c2_;
    int size_ = 0;
};

}  // namespace Utils
}  // namespace pmkd
