// TriangleNode.hpp -- KDNode plus surface area and parent arrays
// (Scene/TriangleNode.h:18-34).
#pragma once
#include "KDNode.hpp"

namespace pmkd {
namespace Scene {

class TriangleNode : public KDNode {
public:
    TriangleNode() = default;
    explicit TriangleNode(const pm_tree &t) : KDNode(t) {}
    float *GetSurfaceAreaData() const { return (float *)t_.surface_area; }
    int *GetParentData() const { return (int *)t_.parent; }
};

}  // namespace Scene
}  // namespace pmkd
