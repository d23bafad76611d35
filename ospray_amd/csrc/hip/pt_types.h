// Device-side data layout of the hip path tracer (plain structs shared by
// host set-up code and kernels).
//
// HBM layout of a committed world:
//   DNode4[]    4-wide BVH nodes, 128 B each (one cache line): the boxes of
//               up to four children, so a traversal step is one 128-byte
//               fetch; TLAS and every BLAS share the array.  child >= 0:
//               inner node; child < 0: leaf, ~child = first << 3 | (count - 1).
//   DPrim[]     48 B per primitive (triangle, quad half or sphere); vertices
//               are stored in object space next to the ids so a leaf is a
//               contiguous run.
//   DInstance[] 128 B: inverse and forward affine transform (columns),
//               BLAS root, model table offset, user id.
//   DModel[]    per geometric model of a group: geometry arrays for
//               post-intersect, material, user id.
//   materials, media, lights, light CDF: small tables.
#pragma once

#include <cstdint>

namespace hpt {

enum PrimKind : uint32_t
{
  PRIM_TRIANGLE = 0,
  PRIM_QUAD0 = 1, // (v0, v1, v3)
  PRIM_QUAD1 = 2, // (v2, v3, v1), reports (1-u, 1-v)
  PRIM_SPHERE = 3
};

// 4-wide BVH node, one 128-byte line: the bounds of up to four children as
// per-axis float4s (child c in lane c), and the child references: >= 0 an
// inner node, < 0 a leaf ~(first << 3 | count - 1).  An unused slot has a
// box at +inf on every axis, which no ray enters.
struct DNode4
{
  float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
  int32_t child[4];
  int32_t pad[4];
};
static_assert(sizeof(DNode4) == 128, "node size");

struct DPrim
{
  // triangles / quad halves: v0, primID | v1, model | kind<<28 | v2, 0
  // spheres: centre, primID | radius,0,0, model | kind<<28 | 0
  float a[3];
  uint32_t primID;
  float b[3];
  uint32_t modelKind;
  float c[3];
  uint32_t pad;
};
static_assert(sizeof(DPrim) == 48, "prim size");

struct DInstance
{
  float rcp[12]; // inverse transform: columns vx, vy, vz, p
  float xfm[12]; // forward transform
  int32_t blasRoot; // node index of the group's BLAS root
  int32_t modelOffset; // first DModel of the group
  uint32_t userID; // kInvalidID if unset
  float epsScale; // max(|xfm.vx.x|, |xfm.vy.y|, |xfm.vz.z|)
  int32_t pad[4];
};
static_assert(sizeof(DInstance) == 128, "instance size");

enum GeomType : int32_t
{
  GEOM_TRIANGLES = 0,
  GEOM_QUADS = 1,
  GEOM_SPHERES = 2
};

struct DModel
{
  const float *vertex; // vec3f
  const uint32_t *index; // vec3ui / vec4ui
  const float *normal; // vec3f or null
  const float *spheres; // vec4f centre + radius
  int32_t geomType;
  int32_t material; // -1: none
  uint32_t userID;
  int32_t invertNormals;
  int32_t opaque; // material transparency is identically zero
  float areaPDF; // as a geometry light (GeometricModel.ispc:51), else 0
  int32_t pad[2];
};

enum MaterialType : int32_t
{
  MAT_OBJ = 0,
  MAT_METAL = 1,
  MAT_GLASS = 2,
  MAT_LUMINOUS = 3
};

struct DMaterial
{
  int32_t type;
  float Kd[3], Tf[3], d; // obj
  float Ks[3], Ns; // obj specular
  float eta[3], k[3], roughness; // metal
  int32_t inside, outside; // glass: indices into the media table
  float radiance[3], transparency; // luminous
  // metal with a spectral 'ior': eta and k at kSpectrumSamples wavelengths,
  // kS directly after etaS (MetalShared.h)
  int32_t spectral;
  float etaS[8], kS[8];
};

struct DMedium
{
  float attenuation[3];
  float ior;
};

enum LightType : int32_t
{
  LIGHT_QUAD = 0,
  LIGHT_DISTANT = 1,
  LIGHT_AMBIENT = 2,
  LIGHT_GEOMETRY = 3 // an emissive model of one instance
};

struct DLight
{
  int32_t type;
  int32_t visible;
  float radiance[3];
  float position[3], edge1[3], edge2[3], nnormal[3];
  float ppdf;
  float frame[9]; // distant: vx, vy, vz
  float irradiance[3];
  float cosAngle, pdf;
  // geometry light: instance, model (DModel index), primitive CDF by area
  // (DScene::geoCDF + geoCdfOffset), 1 / total area
  int32_t geoInst, geoModel, geoNumPrims, geoCdfOffset;
  float geoPdf;
};

// Volumes.  A structured-regular grid of vertex-centred voxels (x fastest)
// at origin + i * spacing, its piecewise-linear transfer function and the
// volumetric model's densityScale / anisotropy / id.  `majorant` holds, per
// macrocell of kMacro^3 cells, densityScale times the transfer function's
// maximum opacity over the macrocell's (min, max) voxel values: the
// majorant of delta tracking, precomputed per volumetric model.
constexpr int kMacro = 16;
constexpr int kTfSubranges = 32;

struct DVolume
{
  const float *voxels;
  const float *majorant; // per macrocell
  const float *tfColor; // vec3f per control point
  const float *tfOpacity;
  int32_t dims[3], mdims[3];
  float origin[3], spacing[3], upper[3];
  int32_t numColors, numOpacities;
  float tfLower, tfUpper;
  float densityScale, anisotropy;
  uint32_t userID;
  int32_t pad;
};

// one (instance, volume) pair: rays meet the volume intervals in the order
// of this list (instance order, then the group's volume order)
struct DVolInst
{
  int32_t inst, vol, local, pad; // instance, DVolume, index within the group
};

// a volume collision in the hit record: inst field = kVolumeHit | pair index
constexpr uint32_t kVolumeHit = 0x40000000u;

struct DScene
{
  const DNode4 *nodes;
  const DPrim *prims;
  const DInstance *instances;
  const DModel *models;
  const DMaterial *materials;
  const DMedium *media;
  // [geometry lights, world lights] and their uniform selection CDF
  const DLight *lights;
  const float *lightsCDF;
  const float *geoCDF; // primitive CDFs of the geometry lights
  const DVolume *volumes;
  const DVolInst *volInsts;
  int32_t tlasRoot;
  int32_t numInstances;
  int32_t numLights;
  int32_t numGeoLights;
  int32_t numMaterials;
  int32_t numVolInsts; // > 0: volumes in the scene (FFO_VOLUME_IN_SCENE)
};

// Path state of one pass, structure of arrays indexed by path slot.
struct DPaths
{
  float *org; // 4 per path: org.xyz, t0
  float *dir; // 4: dir.xyz, (unused)
  float *thr; // 4: throughput.xyz, alpha
  float *contrib; // 4: contribution.xyz, z (primary depth)
  uint32_t *ld; // 4: revIndex, scramble, baseIndexLight, flags
  float *lastP; // 4: last vertex P.xyz, pdf_w
  float *lastNs; // 4: last vertex Ns.xyz, numLightSamples (bits)
  float *hitA; // 4: t, u, v, inst (bits); volume collision: t, 0, 0, kVolumeHit | pair
  float *hitB; // 4: Ng.xyz (object space), model (bits); volume collision: albedo.xyz, 0
  uint32_t *hitPrim; // prim | half << 31
  uint32_t *ids; // 4: primID, geomID, instID, medium
  float *aux; // 16: position.xyz, firstNormal.xyz, normal.xyz, albedo.xyz, pad
  uint32_t *neeN; // light samples of the vertex shaded this bounce (0: none)
  uint32_t *rng; // 4: PCG32 state (lo, hi), stream, (unused): volume sampling
  // NEE vertex record, written by k_shade at a smooth vertex
  float *vtxP; // 4: P.xyz, epsilon
  float *vtxNg; // 4: Ng.xyz, material (bits; volume vertex: -2 - pair)
  float *vtxNs; // 4: Ns.xyz, medium (bits)
  float *vtxThr; // 4: throughput.xyz, light samples (bits)
  float *vtxWo; // 4: wo.xyz, (unused)
  uint32_t *vtxLd; // 4: sampler revIndex, scramble, baseIndexLight, (unused)
  // NEE shadow-ray records, slot = path * maxRec + i
  float *recOrg; // 4: org.xyz, tmax
  float *recDir; // 4: dir.xyz, mis weight
  float *recContrib; // 4: unshaded contribution.xyz, medium (bits)
  float *recResult; // 4: L.xyz, valid flag (with volumes: light contribution
                    // before transmittance and MIS, valid flag)
  float *recVol; // 4: light pdf, BSDF pdf, t0 of the last shadow segment,
                 // (unused): the volume transmittance of k_resolve
  // queues and counters
  uint32_t *queue[2]; // extension queues (ping-pong)
  uint32_t *shadowQueue;
  uint32_t *transQueue; // shadow records that hit only non-opaque surfaces
  uint32_t *neeSampleQueue; // record slots (path * maxRec + sample) to evaluate
  uint32_t *tieQueue; // path slots whose extension hit is re-resolved in FP64
  // queue counters, see kCnt*: the pair (2q, 2q+1) is one 64-bit word so
  // that k_shade appends to extension queue q and to the NEE sample queue
  // with a single atomic
  int32_t *counters;
  uint32_t *overflow; // traversal scratch row per resident thread: stack overflow, recorded hits
  // transparent hits recorded by the occlusion pass for the records of the
  // transparency queue, by queue slot < transCap: 4 float4 and a count
  // (0xffffffff: not recorded, the transparency pass traverses)
  float *transHits;
  uint32_t *transCount;
  int32_t transCap;
  // radiance cache (when enabled): per path slot the vertex record
  // (P, thr.x | wo, thr.y | Ns, thr.z | albedo, -) and the contribution
  // before the vertex; compact queues of the query and training slots
  float *nrcRec; // 16 per path: the query vertex
  float *nrcTrainRec; // 16 per path: the training vertex (self-training)
  float *nrcBefore; // 4 per path
  uint32_t *nrcQueryQueue;
  uint32_t *nrcTrainQueue;
  // node visits, primitive tests, traversals: [0..2] extension rays, [3..5]
  // shadow occlusion pass, [6..8] shadow transparency pass; delta tracking
  // voxel samples, macrocell intervals, tracked segments: [9..11] free
  // paths, [12..14] transmittance estimates
  unsigned long long *stats;
  // cancellation (FrameBuffer::cancelFrame, FrameBuffer.cpp:180-187): the
  // host writes the frame's generation into *cancelFlag (pinned host
  // memory); the per-bounce counter reset reads it and latches
  // *cancelLatch (device memory, one word per framebuffer, cleared when the
  // frame begins), after which the frame's kernels do no more work.  Null
  // in captured frames.
  const uint32_t *cancelFlag;
  uint32_t cancelGen;
  uint32_t *cancelLatch;
};

// queue counters: [2q] entries of extension queue q (q = 0, 1), [2q+1] NEE
// light samples appended while queue q was filled, then:
constexpr int kCntShadow = 4; // shadow-ray records
constexpr int kCntFetchExtend = 5; // work fetch of the extension traversal
constexpr int kCntFetchShadow = 6; // work fetch of the shadow traversal
constexpr int kCntTrans = 7; // records for the transparency pass
constexpr int kCntFetchTrans = 8; // work fetch of the transparency pass
constexpr int kCntTies = 9; // extension rays whose closest hit has a near-tie
constexpr int kNumBounceCounters = 10; // the counters above are reset every bounce
constexpr int kCntNrcQuery = 10; // radiance-cache queries of the pass
constexpr int kCntNrcTrain = 11; // radiance-cache training records of the pass
constexpr int kNumCounters = 12;

// path flags in ld[3]
constexpr uint32_t FLAG_DEPTH_MASK = 0xffu; // bits 0-7
constexpr uint32_t FLAG_SCAT_SHIFT = 8u; // bits 8-15 scattering events
constexpr uint32_t FLAG_FIRST_BOUNCE_LIGHT = 1u << 16;
constexpr uint32_t FLAG_STRAIGHT = 1u << 17;
constexpr uint32_t FLAG_SPEC_TRANS_PATH = 1u << 18;
constexpr uint32_t FLAG_NRC_TRAIN = 1u << 19; // a radiance-cache training path past its record vertex
constexpr uint32_t FLAG_MEDIUM_SHIFT = 20u; // bits 20-31 medium index

struct DFrame
{
  // image and pass
  int32_t width, height;
  int32_t tilesX, tilesY; // 8x8-pixel tasks (the variance estimate's units)
  // path slots: 64-slot tiles of (1 << slotShift) x (64 >> slotShift)
  // pixels, slotTilesX per row, numPixPadded slots per sample
  int32_t slotShift, slotTilesX, numPixPadded;
  int32_t spp, passS0, passCount; // samples [passS0, passS0+passCount)
  uint32_t accumID;
  int32_t targetFrames;
  int32_t projectedDepth;
  uint32_t channels; // OSPFrameBufferChannel mask
  int32_t accumulateVariance;
  int32_t maxRec;
  int32_t residentThreads; // overflow stack rows
  uint32_t idxMask, idxMaskLight; // LDSampler index masks
  // camera (PerspectiveCamera set-up, host side)
  float camOrg[3], camDir00[3], camDu[3], camDv[3], camDirection[3];
  float nearClip;
  // renderer
  int32_t maxDepth, rouletteDepth, maxScatteringEvents;
  uint32_t numFirstLightSamples, numIndirectLightSamples;
  float minContribution, maxRadiance;
  float bgColor[4];
  int32_t filterType; // 0 point, 1 box, 2 LUT
  float filterWidth;
  int32_t lutSize;
  const float *lutCdfX;
  const float *lutCdfY;
  // per-pixel sample accumulators across passes (row-major)
  float *accRGBA; // 4: rgb sum, alpha sum
  float *accAux; // 12: position, firstNormal, normal, albedo sums
  float *accZ;
  uint32_t *accIDs; // 3
  // framebuffer buffers (row-major, null if the channel is absent)
  float *fbColor; // vec4f
  float *fbVariance; // vec4f
  float *fbDepth;
  float *fbNormal; // vec3f
  float *fbAlbedo;
  float *fbFirstNormal;
  float *fbPosition;
  uint32_t *fbPrimID, *fbObjID, *fbInstID;
  float *taskVariance;
  // adaptive accumulation: > 0 renders only the tasks whose taskVariance
  // exceeds it
  float adaptiveThreshold;
  int32_t hasVolumes; // the paths carry a volume sampler (P.rng)
  int32_t stats; // count traversal and delta-tracking work into P.stats
  // radiance cache: paths end at the vertex of depth nrcBounce (-1: off)
  // with a cache query, except every nrcTrainStride-th slot (offset
  // nrcTrainOffset), which runs on and becomes a training record.  With
  // nrcSelfTrain a training path ends one vertex later with a cache query
  // of its own (its target is the radiance gathered up to there plus the
  // cache's estimate beyond); otherwise it runs to its end.
  int32_t nrcBounce;
  uint32_t nrcTrainStride, nrcTrainOffset;
  int32_t nrcSelfTrain;
  // Frames of a captured batch of views on accumulating framebuffers read
  // their frame index from device memory (*frameState, written before each
  // launch of the graph), and derive from it what eager frames get from the
  // host: the variance-frame choice (bit k of varianceCoins is the k-th
  // draw of the framebuffer's mt19937(43), taken on frame 2k), adaptive
  // accumulation (errorThreshold from frame minAdaptiveFrames on) and the
  // odd-frame variance estimate.  Null in eager frames.
  const uint32_t *frameState;
  const uint32_t *varianceCoins;
  float errorThreshold;
  int32_t minAdaptiveFrames;
  // set when the frame was cancelled: the frame's accumulation is skipped
  // (null in captured frames)
  const uint32_t *cancelLatch;
};

constexpr uint32_t kVarianceCoinBits = 1u << 20; // frames up to 2^21 in captured batches

} // namespace hpt
