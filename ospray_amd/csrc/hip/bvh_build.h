// Host-side BVH construction for the hip device: binned SAH over primitive
// boxes, emitting BVH2 child-pair nodes (hpt::DNode, 64 B) in depth-first
// order so a subtree is contiguous in HBM.  The reference gets its BVHs from
// Embree (rtcCommitScene, modules/cpu/common/Group.cpp:129-142,
// World.cpp:167-171; RTC_BUILD_QUALITY_HIGH); the tree shape does not change
// results because traversal resolves ties by primitive key.
#pragma once

#include <cstdint>
#include <vector>

#include "pt_types.h"

namespace hpt {

struct BuildBox
{
  float lo[3], hi[3];
};

struct BuildResult
{
  std::vector<DNode> nodes; // node 0 is the root pair
  std::vector<uint32_t> order; // item ids in leaf order
  BuildBox bounds;
  int maxDepth = 0;
};

// maxLeaf: items per leaf (1..8); node child indices are local to `nodes`
// and leaf `first` indices refer to positions in `order`.
BuildResult buildBVH(const std::vector<BuildBox> &boxes, int maxLeaf, int nthreads);

} // namespace hpt
