// OSPRay C++ wrapper: registers the rkcommon math types (vectors, boxes,
// linear and affine spaces, quaternions) with OSPTypeFor, so they pass to
// setParam and Data directly (ospray_cpp/ext/rkcommon.h of OSPRay 3.x).
// Needs rkcommon's headers on the include path.
#pragma once

#include "rkcommon/math/AffineSpace.h"
#include "rkcommon/math/box.h"

#include "../Traits.h"

// X(rkcommon type, OSPDataType) for every registered type
#define OSPRAY_RKCOMMON_TYPES(X)                                                                                       \
  X(vec2uc, OSP_VEC2UC) X(vec3uc, OSP_VEC3UC) X(vec4uc, OSP_VEC4UC)                                                    \
  X(vec2c, OSP_VEC2C) X(vec3c, OSP_VEC3C) X(vec4c, OSP_VEC4C)                                                          \
  X(vec2us, OSP_VEC2US) X(vec3us, OSP_VEC3US) X(vec4us, OSP_VEC4US)                                                    \
  X(vec2s, OSP_VEC2S) X(vec3s, OSP_VEC3S) X(vec4s, OSP_VEC4S)                                                          \
  X(vec2i, OSP_VEC2I) X(vec3i, OSP_VEC3I) X(vec4i, OSP_VEC4I)                                                          \
  X(vec2ui, OSP_VEC2UI) X(vec3ui, OSP_VEC3UI) X(vec4ui, OSP_VEC4UI)                                                    \
  X(vec2l, OSP_VEC2L) X(vec3l, OSP_VEC3L) X(vec4l, OSP_VEC4L)                                                          \
  X(vec2ul, OSP_VEC2UL) X(vec3ul, OSP_VEC3UL) X(vec4ul, OSP_VEC4UL)                                                    \
  X(vec2f, OSP_VEC2F) X(vec3f, OSP_VEC3F) X(vec4f, OSP_VEC4F)                                                          \
  X(vec2d, OSP_VEC2D) X(vec3d, OSP_VEC3D) X(vec4d, OSP_VEC4D)                                                          \
  X(box1i, OSP_BOX1I) X(box2i, OSP_BOX2I) X(box3i, OSP_BOX3I) X(box4i, OSP_BOX4I)                                      \
  X(box1f, OSP_BOX1F) X(box2f, OSP_BOX2F) X(box3f, OSP_BOX3F) X(box4f, OSP_BOX4F)                                      \
  X(linear2f, OSP_LINEAR2F) X(linear3f, OSP_LINEAR3F)                                                                  \
  X(affine2f, OSP_AFFINE2F) X(affine3f, OSP_AFFINE3F)                                                                  \
  X(quatf, OSP_QUATF)

namespace ospray {

#define OSPRAY_RKCOMMON_REGISTER(t, osp) OSPTYPEFOR_SPECIALIZATION(rkcommon::math::t, osp)
OSPRAY_RKCOMMON_TYPES(OSPRAY_RKCOMMON_REGISTER)
#undef OSPRAY_RKCOMMON_REGISTER

#ifdef OSPRAY_RKCOMMON_DEFINITIONS
#define OSPRAY_RKCOMMON_DEFINE(t, osp) OSPTYPEFOR_DEFINITION(rkcommon::math::t);
OSPRAY_RKCOMMON_TYPES(OSPRAY_RKCOMMON_DEFINE)
#undef OSPRAY_RKCOMMON_DEFINE
#endif

} // namespace ospray
