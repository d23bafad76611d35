// Host-side BVH construction for the hip device: a binned-SAH binary tree
// over primitive boxes, collapsed into 4-wide nodes (hpt::DNode4, 128 B) by
// repeatedly opening the child of largest surface area, emitted in
// depth-first order so a subtree is contiguous in HBM.  The reference gets its BVHs from
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
  std::vector<DNode4> nodes; // node 0 is the root
  std::vector<uint32_t> order; // item ids in leaf order
  BuildBox bounds;
  int maxDepth = 0; // depth of the binary tree
  // traversal stack entries this tree can need: the largest sum of
  // (children - 1) over the inner nodes of a root-to-leaf path
  int stackNeed = 0;
};

// maxLeaf: items per leaf (1..8); node child indices are local to `nodes`
// and leaf `first` indices refer to positions in `order`.
BuildResult buildBVH(const std::vector<BuildBox> &boxes, int maxLeaf, int nthreads);

} // namespace hpt
