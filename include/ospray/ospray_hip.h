/* Extensions of the "hip" device (libospray_module_hip.so).
 *
 * These have no counterpart in the reference C API; they expose what the
 * MI355X device measures so that applications, benchmarks and tests can read
 * it next to the standard API (the reference offers only ospGetTaskDuration,
 * render/RenderingFuture.h:99-103).  All functions act on the current device
 * and return a negative value if it is not a hip device.
 */
#pragma once

#include "ospray.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
  char name[32]; /* kernel / stage name */
  int64_t launches;
  double totalMs; /* sum of hipEvent-timed durations on the device stream */
} OSPHipKernelStat;

/* Per-stage GPU times accumulated while the device parameter
 * "profileKernels" (bool) is set; returns the number of entries written. */
OSPRAY_INTERFACE int ospHipGetKernelStats(OSPHipKernelStat *out, int maxEntries);
OSPRAY_INTERFACE int ospHipResetKernelStats(void);

/* Traversal counters of the last frame rendered while the device parameter
 * "traversalStats" (bool) was set: node visits, primitive tests and
 * traversals of the extension rays (out9[0..2]), of the shadow-ray
 * occlusion pass (out9[3..5]) and of the shadow-ray transparency pass
 * (out9[6..8]). */
OSPRAY_INTERFACE int ospHipGetTraversalStats(uint64_t out9[9]);

/* Renders one frame of each of n views (framebuffer, renderer, camera) of
 * one world in one enqueue and waits for them.  mode 0: eager launches;
 * mode 1: one HIP graph.  The first mode-1 call for a batch renders its
 * frame eagerly (which sizes every buffer) and captures the graph; later
 * calls relaunch it while the batch's objects are unchanged.  Each launch
 * begins the next frame of every view as ospRenderFrame would (the frame
 * index is read from device memory), so accumulating framebuffers
 * accumulate.  viewMs (optional, n floats): GPU milliseconds from the
 * batch start to the end of each view.  Renderers with "nrc" cannot be
 * captured.  Returns 0 on success. */
OSPRAY_INTERFACE int ospHipRenderBatch(int n, const OSPFrameBuffer *framebuffers, const OSPRenderer *renderers,
    const OSPCamera *cameras, OSPWorld world, int mode, float *viewMs);

/* Autotuner of a batch: times the eager batch (best of `reps`) for each
 * candidate resident traversal grid (8, 12, 16 blocks per CU), shading
 * grid (4, 8, 16) and path-state budget (2 GiB, 16 GiB: the samples per
 * pass, i.e. the wavefront queue sizes), then, at the fastest of those,
 * for 1, 2, 4 and 8 lanes (batchStreams), then for the pixel tiles of 64
 * path slots 4x16, 8x8, 16x4 and 32x2 (tileWidth 4, 8, 16, 32), and keeps
 * the fastest as the device's setting.  Returns its batch milliseconds;
 * best5 (5 ints) = traceBlocksPerCU, shadeBlocksPerCU, pathBudgetMB,
 * batchStreams, tileWidth. */
OSPRAY_INTERFACE float ospHipTuneBatch(int n, const OSPFrameBuffer *framebuffers, const OSPRenderer *renderers,
    const OSPCamera *cameras, OSPWorld world, int reps, int *best5);

/* The radiance cache of a "pathtracer" renderer with "nrc" set: the loss of
 * its last training step (relative L2), the number of Adam steps taken so
 * far, and the cache queries and training records of the last frame (all
 * 0 before the first training pass). */
OSPRAY_INTERFACE int ospHipGetNrcState(OSPRenderer renderer, float *loss, int *steps, int64_t *queries,
    int64_t *trainingRecords);

/* Delta-tracking counters of the last frame rendered while "traversalStats"
 * was set: voxel samples (tentative collisions), macrocell intervals and
 * tracked segments of the free-path sampling (out6[0..2]) and of the
 * transmittance estimates (out6[3..5]). */
OSPRAY_INTERFACE int ospHipGetVolumeStats(uint64_t out6[6]);

/* Closest hits of n rays (org.xyz, dir.xyz, t0, t1) against a committed
 * world, traced by the device's BVH kernels.  outTUV: t,u,v; outIDs:
 * instance index, model index within the group, primitive index
 * (0xffffffff on miss); stats2 (optional): node visits, primitive tests. */
OSPRAY_INTERFACE int ospHipTraceRays(
    OSPWorld world, const float *rays8, int n, float *outTUV, uint32_t *outIDs, uint64_t *stats2);

/* The launch-layer view of a committed world (an HptScene, see
 * ospray_hip/hip_pt.h) for callers of the hpt_* entry points; valid until
 * the world is committed again or released.  NULL if the world has not
 * been committed. */
OSPRAY_INTERFACE const void *ospHipGetWorldLaunchScene(OSPWorld world);

/* GPU name, compute-unit count and the BVH statistics of a committed world
 * (total nodes, total primitives, deepest BLAS). */
OSPRAY_INTERFACE int ospHipGetDeviceInfo(char *name, int len, int *numCUs);
OSPRAY_INTERFACE int ospHipGetWorldInfo(OSPWorld world, int64_t out4[4]);

/* BLAS builds of the world's last commit: milliseconds spent building the
 * BLASes it did not find in the device's cache, how many of its BLASes were
 * built on the device (the device parameter "hostBvh" (bool, default
 * true) selects the host binned-SAH builder; false selects the device
 * Morton + PLOC builder), and the traversal stack entries its BVH needs. */
OSPRAY_INTERFACE int ospHipGetWorldBuild(OSPWorld world, double *blasBuildMs, int *deviceBlas, int *stackNeed);

/* Mean duration in ms of `reps` device-to-device copies of `bytes` (rounded
 * down to a multiple of 16) by the device's plain copy kernel: the
 * achievable-bandwidth reference for the roofline figures. */
OSPRAY_INTERFACE int ospHipCopyBandwidth(size_t bytes, int reps, float *msPerCopy);

#ifdef __cplusplus
}
#endif
