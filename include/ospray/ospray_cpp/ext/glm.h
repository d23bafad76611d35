// OSPRay C++ wrapper: registers the glm vector and matrix types with
// OSPTypeFor (ospray_cpp/ext/glm.h of OSPRay 3.x).  glm's column-major
// mat3x3 / mat2x2 match OSP_LINEAR3F / OSP_LINEAR2F (columns vx, vy, vz),
// and mat4x3 / mat3x2 (4 or 3 columns) match OSP_AFFINE3F / OSP_AFFINE2F
// (linear part, then translation).  Needs glm's headers on the include path.
#pragma once

#include "glm/mat2x2.hpp"
#include "glm/mat3x2.hpp"
#include "glm/mat3x3.hpp"
#include "glm/mat4x3.hpp"
#include "glm/vec2.hpp"
#include "glm/vec3.hpp"
#include "glm/vec4.hpp"

#include "../Traits.h"

#define OSPRAY_GLM_TYPES(X)                                                                                            \
  X(vec2, OSP_VEC2F) X(vec3, OSP_VEC3F) X(vec4, OSP_VEC4F)                                                             \
  X(ivec2, OSP_VEC2I) X(ivec3, OSP_VEC3I) X(ivec4, OSP_VEC4I)                                                          \
  X(uvec2, OSP_VEC2UI) X(uvec3, OSP_VEC3UI) X(uvec4, OSP_VEC4UI)                                                       \
  X(mat2x2, OSP_LINEAR2F) X(mat3x3, OSP_LINEAR3F) X(mat3x2, OSP_AFFINE2F) X(mat4x3, OSP_AFFINE3F)

namespace ospray {

#define OSPRAY_GLM_REGISTER(t, osp) OSPTYPEFOR_SPECIALIZATION(glm::t, osp)
OSPRAY_GLM_TYPES(OSPRAY_GLM_REGISTER)
#undef OSPRAY_GLM_REGISTER

#ifdef OSPRAY_GLM_DEFINITIONS
#define OSPRAY_GLM_DEFINE(t, osp) OSPTYPEFOR_DEFINITION(glm::t);
OSPRAY_GLM_TYPES(OSPRAY_GLM_DEFINE)
#undef OSPRAY_GLM_DEFINE
#endif

} // namespace ospray
