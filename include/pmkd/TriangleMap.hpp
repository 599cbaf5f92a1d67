// TriangleMap.hpp -- the kd-tree over a GeometryList
// (Utils/CUDA/TriangleMap.h:36-82): Create() runs the upper creator and
// the lower creator (TriangleMap.cu:51-62); the setters are the
// reference's options.  Owns the pm_ctx that holds the tree.
#pragma once
#include <ostream>

#include "GeometryList.hpp"
#include "TriangleNode.hpp"

namespace pmkd {
namespace Utils {

class TriangleMap {
public:
    enum LowerAlgorithm { BITMAP = PM_LOWER_BITMAP, BALANCED = PM_LOWER_BALANCED, SAH = PM_LOWER_SAH };
    enum SplitMethod { BOX = PM_SPLIT_BOX, DIVIDE = PM_SPLIT_DIVIDE, SPLIT = PM_SPLIT_SPLIT };

    explicit TriangleMap(GeometryList *geom) : geom_(geom) {
        if (pm_create(&ctx_, 0) != PM_OK) throw Exception(PM_ERR_HIP, "no HIP device");
        pm_build_params_default(&params_);
        hip_ok(hipEventCreate(&e0_), "hipEventCreate");
        hip_ok(hipEventCreate(&e1_), "hipEventCreate");
    }
    ~TriangleMap() {
        if (ctx_) pm_destroy(ctx_);
        (void)hipEventDestroy(e0_);
        (void)hipEventDestroy(e1_);
    }
    TriangleMap(const TriangleMap &) = delete;
    TriangleMap &operator=(const TriangleMap &) = delete;

    void Create(hipStream_t stream = nullptr) {
        hip_ok(hipEventRecord(e0_, stream), "hipEventRecord");
        check(ctx_, pm_build(ctx_, (const float *)geom_->GetP0Data(), (const float *)geom_->GetP1Data(),
                             (const float *)geom_->GetP2Data(), geom_->GetSize(), &params_, stream));
        hip_ok(hipEventRecord(e1_, stream), "hipEventRecord");
        hip_ok(hipEventSynchronize(e1_), "hipEventSynchronize");
        hip_ok(hipEventElapsedTime(&constructionTime_, e0_, e1_), "hipEventElapsedTime");
    }

    LowerAlgorithm GetLowerAlgorithm() const { return (LowerAlgorithm)params_.lower_algorithm; }
    void SetLowerAlgorithm(LowerAlgorithm l) { params_.lower_algorithm = l; }
    float GetConstructionTime() const { return constructionTime_; }
    void SplitEmptySpace(bool s) { params_.empty_space_splitting = s; }
    bool IsSplittingEmptySpace() const { return params_.empty_space_splitting != 0; }
    void SetSplitMethod(SplitMethod s) { params_.split_method = s; }
    SplitMethod GetSplitMethod() const { return (SplitMethod)params_.split_method; }
    void SetPropagateBoundingBox(bool p) { params_.propagate_aabbs = p; }
    bool GetPropagateBoundingBox() const { return params_.propagate_aabbs != 0; }
    void SetTraversalCost(float t) { params_.traversal_cost = t; }
    float GetTraversalCost() const { return params_.traversal_cost; }
    // TriangleMapUpperCreator::SetEmptySpaceThreshold (TriangleMapUpperCreator.h:86)
    void SetEmptySpaceThreshold(float t) { params_.empty_space_threshold = t; }
    // KDNODE_MAX_LOWER_SIZE (KDNode.h:21), a compile-time constant there
    void SetLowerSize(int n) { params_.lower_size = n; }
    pm_build_params &Params() { return params_; }

    GeometryList *GetGeometry() const { return geom_; }
    Scene::TriangleNode GetNodes() const {
        pm_tree t;
        check(ctx_, pm_tree_view(ctx_, &t));
        return Scene::TriangleNode(t);
    }
    int *GetPrimitiveIndices() const { return (int *)View().prim_indices; }
    int GetPrimitiveIndicesSize() const { return View().num_prim_indices; }
    pm_tree View() const {
        pm_tree t;
        check(ctx_, pm_tree_view(ctx_, &t));
        return t;
    }
    pm_stats Stats(hipStream_t stream = nullptr) const {
        pm_stats s;
        check(ctx_, pm_tree_stats(ctx_, &s, stream));
        return s;
    }
    // PrintTree (TriangleMap.cu:110-116)
    void PrintTree(std::ostream &os) const {
        const pm_stats s = Stats();
        os << "Start primitives: " << geom_->GetSize() << "\n"
           << "End primitives: " << s.num_prim_indices << "\n"
           << "Tree nodes: " << s.num_nodes << " (leaves " << s.leaves << ", max depth " << s.max_depth << ")\n";
    }
    pm_ctx *Context() const { return ctx_; }

private:
    GeometryList *geom_;
    pm_ctx *ctx_ = nullptr;
    pm_build_params params_;
    hipEvent_t e0_ = nullptr, e1_ = nullptr;
    float constructionTime_ = 0.0f;
};

}  // namespace Utils
}  // namespace pmkd
