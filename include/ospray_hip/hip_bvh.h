/* Device BVH build of the hip path tracer (libospray_module_hip.so).
 *
 * Replaces the BLAS build the reference hands to Embree when a group is
 * committed (rtcNewScene / rtcAttachGeometry / rtcCommitScene,
 * modules/cpu/common/Group.cpp:93-143, RTC_BUILD_QUALITY_HIGH): one BVH over
 * every primitive of a list of geometries (triangles, quads as two
 * triangles, spheres), built on the GPU and written in the traversal
 * kernels' layout (hpt::DNode4 4-wide nodes, hpt::DPrim records in leaf
 * order; ospray_amd/csrc/hip/pt_types.h).
 *
 * Method: primitive boxes -> 30-bit Morton codes of the box centres on a
 * cubic grid over the centre bounds -> stable LSD radix sort -> PLOC
 * agglomerative clustering (Meister and Bittner 2018: every cluster picks
 * the neighbour within a window of the sorted order whose merged box has
 * the smallest surface area; mutual pairs merge) -> binary tree -> 4-wide nodes, each opening
 * its child of largest surface area until it holds four, subtrees of at
 * most two primitives becoming leaves.
 *
 * hpt_bvh_build runs synchronously on the given stream (it reads back the
 * cluster count between clustering steps through hostWords), never
 * allocates and never throws.  The caller provides the workspace
 * (hpt_bvh_workspace_bytes), the output arrays (capacity: one node and one
 * primitive record per item) and hostWords, 4 uint64 words of pinned host
 * memory (hipHostMalloc).
 */
#pragma once

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef HPT_API
#define HPT_API __attribute__((visibility("default")))
#endif

/* One geometry of the BVH: device arrays as uploaded at geometry commit.
 * geomType: 0 triangles (vec3ui index), 1 quads (vec4ui index, two items
 * per quad: (v0,v1,v3) and (v2,v3,v1)), 2 spheres (vec4f centre+radius). */
typedef struct HptBvhGeometry
{
  const float *vertex;
  const uint32_t *index;
  const float *spheres;
  int32_t geomType;
  uint32_t numPrims;
} HptBvhGeometry;

typedef struct HptBvhResult
{
  uint32_t numNodes; /* 4-wide nodes written, root first */
  uint32_t numItems; /* primitive records written */
  int32_t stackNeed; /* largest sum of (children - 1) over a root-to-leaf path */
  int32_t levels; /* levels of the 4-wide tree */
  int32_t clusterSteps; /* clustering iterations */
  float lo[3], hi[3]; /* bounds of every primitive */
} HptBvhResult;

/* Items (primitive records) of the geometries: quads count twice. */
HPT_API uint64_t hpt_bvh_items(const HptBvhGeometry *geoms, int numGeoms);
/* Bytes of workspace hpt_bvh_build needs for numItems items. */
HPT_API size_t hpt_bvh_workspace_bytes(uint64_t numItems);
/* geoms: host array of numGeoms entries (copied into the workspace).
 * radius: the clustering window, 1..32 clusters on either side (16 is the
 * usual choice; larger windows build better trees more slowly).
 * Requires 3 <= items < 2^28.  outNodes: capacity numItems nodes;
 * outPrims: numItems records.  Node child references are local to
 * outNodes, leaf `first` indices local to outPrims. */
HPT_API hipError_t hpt_bvh_build(const HptBvhGeometry *geoms, int numGeoms, int radius, void *workspace,
    size_t workspaceBytes, void *outNodes, void *outPrims, uint64_t *hostWords, HptBvhResult *result, hipStream_t s);
/* dst[i] = src[i] for n nodes, inner child references + nodeOffset, leaf
 * first indices + primOffset (placing a BLAS in the world's arrays). */
HPT_API hipError_t hpt_bvh_relocate(void *dst, const void *src, size_t n, int32_t nodeOffset, uint32_t primOffset,
    hipStream_t s);

#ifdef __cplusplus
}
#endif
