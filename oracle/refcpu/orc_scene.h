// ORACLE — test infrastructure only.
//
// Flat, C-ABI scene description consumed by the oracle renderer
// (liborc.so).  It mirrors the OSPRay object graph that the path tracer sees
// after commit: geometries -> geometric models -> groups -> instances ->
// world, plus materials, lights, camera, renderer and framebuffer parameters
// (semantics and defaults of modules/cpu/{geometry,common,render,lights,
// camera,fb}/*.cpp, cited where each field is consumed in orc_render.cpp).
#pragma once

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum OrcGeomType
{
  ORC_GEOM_TRIANGLES = 0,
  ORC_GEOM_QUADS = 1,
  ORC_GEOM_SPHERES = 2
};

typedef struct
{
  int32_t type; // OrcGeomType
  const float *vertex; // vec3f positions (meshes)
  int64_t numVertices;
  const uint32_t *index; // vec3ui / vec4ui per primitive (meshes)
  int64_t numPrimitives;
  const float *normal; // optional vec3f per-vertex normals (meshes)
  const float *spheres; // vec4f centre+radius per sphere (spheres)
} OrcGeometry;

typedef struct
{
  int32_t geometry; // index into geometries
  int32_t material; // index into materials, -1 = none
  uint32_t userID; // 0xffffffff = not set ("id" parameter)
  int32_t invertNormals;
} OrcModel;

typedef struct
{
  int32_t firstModel; // range into groupModels
  int32_t numModels;
  int32_t firstVolume; // range into groupVolumes
  int32_t numVolumes;
} OrcGroup;

// a volumetric model: structured-regular volume (vertex-centred voxels,
// x fastest), piecewise-linear transfer function, densityScale, anisotropy
// (VolumetricModel.cpp:74-82, LinearTransferFunction.cpp:18-28)
typedef struct
{
  const float *voxels;
  int64_t dims[3];
  float gridOrigin[3], gridSpacing[3];
  const float *tfColor; // vec3f
  int32_t numColors;
  const float *tfOpacity;
  int32_t numOpacities;
  float tfValueRange[2];
  float densityScale, anisotropy;
  uint32_t userID; // 0xffffffff = not set
} OrcVolume;

typedef struct
{
  int32_t group;
  uint32_t userID; // 0xffffffff = not set
  float xfm[12]; // affine3f: linear columns vx, vy, vz, then p
} OrcInstance;

enum OrcMaterialType
{
  ORC_MAT_OBJ = 0,
  ORC_MAT_METAL = 1,
  ORC_MAT_GLASS = 2,
  ORC_MAT_LUMINOUS = 3
};

typedef struct
{
  int32_t type;
  float kd[3], d, tf[3]; // obj
  float eta[3], k[3], roughness; // metal
  float etaInside, etaOutside; // glass
  float attenuationColorInside[3], attenuationColorOutside[3];
  float attenuationDistance;
  float color[3], intensity, transparency; // luminous
  float ks[3], ns; // obj specular
  const float *ior; // metal: (wavelength nm, eta, k) triples ordered by wavelength, or null
  int64_t numIor;
} OrcMaterial;

enum OrcLightType
{
  ORC_LIGHT_QUAD = 0,
  ORC_LIGHT_DISTANT = 1,
  ORC_LIGHT_AMBIENT = 2
};

typedef struct
{
  int32_t type;
  int32_t visible;
  float color[3];
  float intensity;
  int32_t intensityQuantity; // OSPIntensityQuantity, -1 = light default
  float position[3], edge1[3], edge2[3]; // quad
  float direction[3], angularDiameter; // distant
} OrcLight;

typedef struct
{
  const OrcGeometry *geometries;
  int32_t numGeometries;
  const OrcModel *models;
  int32_t numModels;
  const int32_t *groupModels; // model indices, grouped by OrcGroup ranges
  const OrcGroup *groups;
  int32_t numGroups;
  const OrcInstance *instances;
  int32_t numInstances;
  const OrcMaterial *materials;
  int32_t numMaterials;
  const OrcLight *lights;
  int32_t numLights;
  const OrcVolume *volumes;
  int32_t numVolumes;
  const int32_t *groupVolumes; // volume indices, grouped by OrcGroup ranges
} OrcScene;

typedef struct
{
  float position[3], direction[3], up[3];
  float fovy, aspect, nearClip;
} OrcCamera;

typedef struct
{
  int32_t pixelSamples;
  int32_t maxPathLength;
  int32_t roulettePathLength;
  int32_t maxScatteringEvents;
  int32_t lightSamples;
  int32_t limitIndirectLightSamples;
  float minContribution;
  float maxContribution;
  float backgroundColor[4];
  int32_t pixelFilter; // OSPPixelFilterType
  float varianceThreshold;
  int32_t geometryLights; // emissive geometry as NEE lights (PathTracer.cpp:49)
} OrcRenderer;

#ifdef __cplusplus
}
#endif
