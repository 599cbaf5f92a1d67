// pmkd.hpp -- the C++ API with the reference's class names, over the C ABI
// of pmkd.h.  Link with -lpmkd -lamdhip64.
#pragma once
#include "BruteTracer.hpp"
#include "CUDAMeshNode.hpp"
#include "GeometryList.hpp"
#include "IRayTracer.hpp"
#include "KDNode.hpp"
#include "PhotonTracer.hpp"
#include "RayTracer.hpp"
#include "ShortStack.hpp"
#include "TriangleMap.hpp"
#include "TriangleNode.hpp"
