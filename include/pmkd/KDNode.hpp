// KDNode.hpp -- read-only view of the tree arrays, with the getters of
// Scene/KDNode.h:85-94 (device pointers, reference SoA layout).
#pragma once
#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>

#include "../pmkd.h"

namespace pmkd {
namespace Scene {

class KDNode {
public:
    static const int MAX_LOWER_SIZE = 64;        // KDNode.h:21 (64-bit bitmap branch)
    typedef long long bitmap;                    // KDNode.h:37-41
    typedef bitmap amount;
    static const char LEAF = 0, X = 1, Y = 2, Z = 3;   // KDNode.h:62-65

    KDNode() = default;
    explicit KDNode(const pm_tree &t) : t_(t) {}

    char *GetInfoData() const { return (char *)t_.info; }
    float *GetSplitPositionData() const { return (float *)t_.split_pos; }
    float4 *GetAabbMinData() const { return (float4 *)t_.aabb_min; }
    float4 *GetAabbMaxData() const { return (float4 *)t_.aabb_max; }
    int *GetPrimitiveIndexData() const { return (int *)t_.primitive_index; }
    amount *GetPrimitiveAmountData() const { return (amount *)t_.primitive_bitmap; }
    bitmap *GetPrimitiveBitmapData() const { return (bitmap *)t_.primitive_bitmap; }
    int2 *GetChildrenData() const { return (int2 *)t_.children; }
    int GetSize() const { return t_.num_nodes; }

protected:
    pm_tree t_{};
};

}  // namespace Scene
}  // namespace pmkd
